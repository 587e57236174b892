"""Benchmark: SIMPLE iterations per second of the flow-solver step on one MI355X.

    python bench.py [--gpus 1] [--steps 50] [--warmup 3] [--m 128]

A step is one SIMPLE iteration (one pass of FlowSolver::timestep's loop body,
FlowSolver.inl:3190-3267) on the box mesh of modifier m (2m x m x m hex cells;
m = 128 -> 4,194,304 cells), in simple mode (pressure-correction system
re-assembled and solved to 1e-8 relative residual every iteration), with every
input already resident in HBM.  Each pressure-solve form is timed on its own
context: FP64 PCG-AMG; FP32 PCG-AMG inside FP64 iterative refinement (the
fine-level smoothing sweeps on an FP16 copy of D^-1 A); and that mixed form
with the learned pressure-correction guess.  `value` is the iterations/s of
the learned + mixed form, the form the metric names (`speedup_vs_fp64` beside
it); the other forms' rates are listed under `forms`.

A profiled pass (before the timed one) brackets every kernel class with HIP
events on the context's stream (no synchronisation) and divides its
algorithmic bytes by its time; the same pass times the iteration's stages
(gradients, velocity, pressure, turbulence and scalar assembly and solves:
`stage_ms_per_step`).  The dominant single-kernel class is then timed
again with HIP events inside the timed region itself: `roofline` is that
kernel's algorithmic bytes per launch over its average launch duration,
against the 8 TB/s HBM3E peak; `traffic` is its HBM bytes per launch from the
FETCH_SIZE/WRITE_SIZE counter summary in profiles/ (null when absent), and
the rates of both byte counts are reported against the peak and against the
on-box copy probe.  Algorithmic bytes count each array a kernel touches once
per launch; a class whose algorithmic rate comes out above the HBM peak is
refused (listed under `refused_classes`, never the roofline kernel): its
byte count or its cache residency is wrong.  `cpu_baseline` is the CPU oracle
(one core) timed on a bounded sample of the same workload (3 SIMPLE
iterations, BiCGStab started as the GPU forms start it), run in a child
process after the GPU forms are measured.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SIMPLE iter/s at 4.2M cells: learned+mixed vs FP64 GPU at equal residual"
FIELDS = [("U", 0), ("V", 1), ("W", 2), ("P", 3), ("TE", 5), ("ED", 6), ("TEM", 8), ("FUL", 9), ("PRO", 10),
          ("VARF", 11), ("VARP", 12)]
HBM_PEAK_GBS = 8000.0


# Loop-1 pressure solves of the learned form start from the Galerkin projection
# onto W = [net guess, last 4 loop-1 solutions] (deflation space); the
# recycled form is the same without the net (what the net adds on its own).
RECYCLE = 4


def run_form(mc, m, form, steps, warmup, profile_steps, roof_class=None, extra=None, mode="simple"):
    psolve = mc.PSOLVE_MIXED if "mixed" in form else mc.PSOLVE_FP64
    use_net = form.startswith("learned")
    recycle = RECYCLE if (use_net or form.startswith("recycled")) else 0
    opts = dict(mode=mc.MODE_SIMPLE if mode == "simple" else mc.MODE_COMPAT, psolve=psolve, p_rtol=1e-8,
                mom_rtol=1e-5, learned=1 if use_net else 0, p_recycle=recycle)
    opts.update(extra or {})
    ctx = mc.Context(m, **opts)
    if use_net:
        from minicombust_app_amd import learned
        ctx.mpnn_load(learned.load_weights())
    it = 0
    for _ in range(warmup):
        ctx.simple_iteration(it % 3)
        it += 1
    ctx.sync()
    s0 = ctx.stats()   # the AMG hierarchy was built during the warm-up
    amg_setup = {"builds": s0["amg_setups"], "ms": round(s0["amg_setup_ms"], 2)}
    # profiled pass: every kernel class bracketed by HIP events (no syncs)
    kern = {}
    stages = None
    if profile_steps > 0:
        ctx.sync()
        ctx.profile(True)
        ctx.profile_reset()
        ctx.stage_enable(True)
        ctx.stage_reset()
        for _ in range(profile_steps):
            ctx.simple_iteration(it % 3)
            it += 1
        prof = ctx.profile_read()
        stages = {k: round(v["ms"] / profile_steps, 4) for k, v in ctx.stage_read().items()}
        ctx.stage_enable(False)
        ctx.profile(False)
        for name, v in prof.items():
            if v["calls"] == 0:
                continue
            ms_avg = v["ms"] / v["calls"]
            rate = (v["bytes"] / v["calls"]) / (ms_avg * 1e-3) if v["bytes"] > 0 and ms_avg > 0 else None
            kern[name] = {"ms_per_step": v["ms"] / profile_steps, "calls_per_step": v["calls"] / profile_steps,
                          "avg_launch_ms": ms_avg, "bytes_per_launch": v["bytes"] / v["calls"]}
            if name.startswith("mpnn"):   # these classes count MFMA flops, not bytes
                kern[name]["achieved_TFLOPs"] = rate / 1e12 if rate else None
                kern[name]["achieved_GBs"] = None
            else:
                kern[name]["achieved_GBs"] = rate / 1e9 if rate else None
    refused = {k: v["achieved_GBs"] for k, v in kern.items()
               if v.get("achieved_GBs") and v["achieved_GBs"] > HBM_PEAK_GBS}
    for k in refused:
        print("# class %s refused: algorithmic rate %.0f GB/s above the HBM peak" % (k, refused[k]),
              file=sys.stderr, flush=True)
        kern[k]["refused"] = "algorithmic rate above the HBM peak"
    if roof_class is None and kern:
        cands = [(v["ms_per_step"], k) for k, v in kern.items()
                 if v["achieved_GBs"] and k in mc.KC_KERNELS and k not in refused]
        roof_class = max(cands)[1] if cands else None
    # timed pass: only the roofline kernel's launches carry events
    ctx.sync()
    ctx.profile(roof_class is not None, classes=[roof_class] if roof_class else None)
    ctx.profile_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.simple_iteration(it % 3)
        it += 1
    ctx.sync()
    t1 = time.perf_counter()
    st = ctx.stats()
    live = None
    if roof_class is not None:
        v = ctx.profile_read()[roof_class]
        if v["calls"]:
            live = {"kernel_class": roof_class, "kernel": mc.KC_KERNELS.get(roof_class),
                    "avg_launch_ms": v["ms"] / v["calls"], "launches": v["calls"],
                    "bytes_per_launch": v["bytes"] / v["calls"]}
    ctx.profile(False)
    out = {
        "iter_per_s": steps / (t1 - t0),
        "ms_per_step": 1000.0 * (t1 - t0) / steps,
        "pressure_ms_per_step": st["p_solve_ms"] / steps,
        "pressure_iters_per_step": st["p_iters"] / steps,
        "pressure_solves_per_step": st["p_solves"] / steps,
        "refinement_steps_per_step": st["p_outer"] / steps,
        "last_rel_residual": st["p_relres_max"],
        "momentum_iters_per_step": st["mom_iters"] / steps,
        "learned_guesses_per_step": st["p_guesses"] / steps,
        "learned_guess_residual_ratio": (st["p_guess_ratio_sum"] / st["p_guesses"]) if st["p_guesses"] else None,
        "deflation_starts_per_step": st["p_defl_starts"] / steps,
        "deflation_start_residual_ratio": (st["p_defl_ratio_sum"] / st["p_defl_starts"]) if st["p_defl_starts"] else None,
        "deflation_vectors_per_start": (st["p_defl_vectors"] / st["p_defl_starts"]) if st["p_defl_starts"] else None,
        "scalar_iters_per_step": st["scalar_iters"] / steps,
        "kernels": kern,
        "refused_classes": refused or None,
        "roof_live": live,
        # final fields after the same iteration sequence in every form
        "fields": {name: ctx.phi(f)[:ctx.N].copy() for name, f in FIELDS},
        "memory": ctx.memory(),
        "amg_half": bool(ctx.opts.amg_half) and psolve == mc.PSOLVE_MIXED,
        "amg_setup": amg_setup,
        "stage_ms_per_step": stages,
    }
    if use_net and "mpnn_fwd" not in kern:
        # the measured policy can leave the guess off through the profiled steps:
        # time the net's forward pass on one pressure system recorded after them
        ctx.mpnn_collect(1)
        ctx.simple_iteration(it % 3)
        ctx.mpnn_collect(0)
        if ctx.mpnn_dataset_size() > 0:
            ctx.sync()
            ctx.profile(True, classes=["mpnn_fwd"])
            ctx.profile_reset()
            for _ in range(3):
                ctx.mpnn_forward_pair(0)
            v = ctx.profile_read()["mpnn_fwd"]
            ctx.profile(False)
            if v["calls"]:
                ms_avg = v["ms"] / v["calls"]
                kern["mpnn_fwd"] = {"ms_per_step": None, "calls_per_step": None, "avg_launch_ms": ms_avg,
                                    "bytes_per_launch": v["bytes"] / v["calls"], "achieved_GBs": None,
                                    "achieved_TFLOPs": (v["bytes"] / v["calls"]) / (ms_avg * 1e-3) / 1e12}
    ctx.close()
    return out


def mfma_of(forms):
    """MFMA throughput of the learned guess's forward pass (BF16 dense peak)."""
    for name, v in forms.items():
        k = v["kernels"].get("mpnn_fwd")
        if k and k["avg_launch_ms"] > 0:
            tf = k["bytes_per_launch"] / (k["avg_launch_ms"] * 1e-3) / 1e12
            return {"form": name, "flops_per_forward": k["bytes_per_launch"], "avg_forward_ms": round(k["avg_launch_ms"], 4),
                    "achieved": round(tf, 2), "peak": 2500.0, "unit": "TFLOP/s", "frac": round(tf / 2500.0, 5)}
    return None


def field_errors(forms):
    """max |x - x_fp64| / max |x_fp64| per field, for each form against fp64."""
    import numpy as np
    if "fp64" not in forms:
        return None
    ref = forms["fp64"]["fields"]
    out = {}
    for name, v in forms.items():
        if name == "fp64":
            continue
        out[name] = {f: float(np.max(np.abs(v["fields"][f] - ref[f])) / max(np.max(np.abs(ref[f])), 1e-300))
                     for f in ref}
    return out


def traffic_of(kernel, m):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE/WRITE_SIZE
    summary (rocprofv3 --pmc runs, scaled by a plain copy of known size)."""
    path = os.path.join(ROOT, "profiles", "traffic_m%d.txt" % m)
    if not kernel or not os.path.exists(path):
        return None

    def norm(name):
        return name.split("(")[0].replace("void", "").replace(" ", "")

    for ln in open(path):
        f = ln.split()
        if len(f) >= 2 and not ln.startswith("#") and norm(f[0]) == norm(kernel):
            try:
                return float(f[-1])
            except ValueError:
                return None
    return None


def roofline_of(live, m, probe):
    """achieved = algorithmic bytes per launch / average launch time; the
    measured traffic's rate beside it, each against the HBM peak and the
    copy probe's rate."""
    if not live:
        return None
    sec = live["avg_launch_ms"] * 1e-3
    gbs = live["bytes_per_launch"] / sec / 1e9
    traffic = traffic_of(live["kernel"], m)
    tgbs = traffic / sec / 1e9 if traffic else None
    pr = probe["GBs"] if probe else None
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": live["kernel"], "kernel_class": live["kernel_class"],
            "bytes_per_launch": live["bytes_per_launch"], "avg_launch_ms": round(live["avg_launch_ms"], 5),
            "launches_timed": live["launches"],
            "frac_of_copy_probe": round(gbs / pr, 4) if pr else None,
            "traffic_over_algorithmic": round(traffic / live["bytes_per_launch"], 3) if traffic else None,
            "traffic_GBs": round(tgbs, 1) if tgbs else None,
            "traffic_frac": round(tgbs / HBM_PEAK_GBS, 4) if tgbs else None,
            "traffic_frac_of_copy_probe": round(tgbs / pr, 4) if (tgbs and pr) else None}


def copy_probe(mc, mb, reps):
    """A plain device copy of known size: the on-box HBM bandwidth probe and
    the known-size kernel (mc::k_copy_probe) that scales the traffic counters."""
    ctx = mc.Context(4, mode=mc.MODE_SIMPLE)
    nbytes = mb << 20
    ms = ctx.copy_probe(nbytes, reps)
    ctx.close()
    gbs = 2.0 * nbytes / (ms * 1e-3) / 1e9
    return {"kernel": "mc::k_copy_probe", "bytes_read": nbytes, "bytes_written": nbytes, "reps": reps,
            "avg_launch_ms": round(ms, 5), "GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--m", type=int, default=128)
    ap.add_argument("--forms", default="fp64,mixed,recycled_mixed,learned_mixed")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--cpu-sample-m", type=int, default=128)
    ap.add_argument("--cpu-iters", type=int, default=3, help="timed SIMPLE iterations of the CPU baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--roofline-class", default=None, help="kernel class for the roofline (default: dominant)")
    ap.add_argument("--copy-probe-mb", type=int, default=1024, help="size of the HBM copy probe (MiB)")
    ap.add_argument("--mode", choices=["simple", "compat"], default="simple",
                    help="pressure-correction semantics (compat: the reference CPU path as written)")
    ap.add_argument("--opt", action="append", default=[], metavar="KEY=VALUE",
                    help="override an mc_options field for every form (repeatable)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from minicombust_app_amd import mcflow as mc
    from minicombust_app_amd import learned
    forms = {}
    skipped = {}
    for form in args.forms.split(","):
        if form.startswith("learned") and not os.path.exists(learned.WEIGHTS):
            skipped[form] = "no trained weights (%s)" % os.path.relpath(learned.WEIGHTS, ROOT)
            print("# %s skipped: %s" % (form, skipped[form]), file=sys.stderr, flush=True)
            continue
        extra = {}
        for kv in args.opt:
            k, v = kv.split("=", 1)
            extra[k] = float(v) if "." in v or "e" in v else int(v)
        forms[form] = run_form(mc, args.m, form, args.steps, args.warmup, args.profile_steps, args.roofline_class,
                               extra, args.mode)
        print("# %s: %.2f SIMPLE iter/s, pressure %.2f ms/iter (%.1f PCG iters), rel.res %.2e" % (
            form, forms[form]["iter_per_s"], forms[form]["pressure_ms_per_step"],
            forms[form]["pressure_iters_per_step"], forms[form]["last_rel_residual"]), file=sys.stderr, flush=True)

    # the library's warm_start default (-1) starts the solves from the current fields in simple mode
    ws = [int(kv.split("=", 1)[1]) for kv in args.opt if kv.split("=", 1)[0] == "warm_start"]
    warm = (ws[-1] > 0 or (ws[-1] < 0 and args.mode == "simple")) if ws else args.mode == "simple"
    probe = copy_probe(mc, args.copy_probe_mb, 10) if args.copy_probe_mb > 0 else None

    # the metric's form: learned + mixed; without it (weights missing or the
    # form not asked for) the fastest form that met the residual, named in config
    ok = {k: v for k, v in forms.items() if v["last_rel_residual"] <= 1e-8}
    head = "learned_mixed" if "learned_mixed" in forms else max(ok or forms, key=lambda k: forms[k]["iter_per_s"])
    h = forms[head]

    cpu = None
    if not args.no_cpu_baseline:
        print("# cpu baseline: oracle at m=%d" % args.cpu_sample_m, file=sys.stderr, flush=True)
        try:
            out = subprocess.run([sys.executable, "-m", "oracle.cpu_bench", "--m", str(args.cpu_sample_m),
                                  "--iters", str(args.cpu_iters)], cwd=ROOT, capture_output=True, text=True,
                                 timeout=900)
            cpu = json.loads(out.stdout.strip().splitlines()[-1])
        except Exception as e:  # report, never fake
            cpu = {"error": "cpu baseline failed: %s" % e}

    N = 2 * args.m ** 3
    half = h["amg_half"]
    dtype = {"fp64": "f64", "mixed": "f64+f32" + ("+f16" if half else ""),
             "recycled_mixed": "f64+f32" + ("+f16" if half else ""),
             "learned_mixed": "f64+f32" + ("+f16" if half else "") + "+bf16", "learned_fp64": "f64+bf16"}[head]
    precision = {"fp64": "FP64 everywhere",
                 "mixed": "FP64 refinement residual, update and coarsest solve; FP32 PCG, SpMV and AMG levels"
                          + ("; fine-level smoothing sweeps read an FP16 copy of D^-1 A" if half else ""),
                 "learned_fp64": "FP64 solve; net BF16 weights and MFMA operands, FP32 accumulation"}
    precision["recycled_mixed"] = precision["mixed"] + "; FP64 deflation-space start"
    precision["learned_mixed"] = (precision["mixed"] + "; FP64 deflation-space start; net BF16 weights and MFMA "
                                  "operands, FP32 accumulation")
    line = {
        "metric": METRIC,
        "value": round(h["iter_per_s"], 3),
        "unit": "SIMPLE iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(h["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (reference box mesh and default initial/boundary fields)",
        "config": {"workload": "box mesh modifier %d (%d hex cells), %s mode, pressure rel. residual 1e-8, "
                               "momentum/scalar BiCGStab rtol 1e-5 (||r|| <= rtol ||b||, %s)"
                               % (args.m, N, args.mode, "started from the current fields" if warm else "started from zero"),
                   "modifier": args.m, "cells": N, "form": head, "precision": precision[head],
                   "pressure_start": {"fp64": "zero", "mixed": "zero",
                                      "recycled_mixed": "Galerkin projection onto the last %d loop-1 solutions"
                                                        % RECYCLE,
                                      "learned_mixed": "Galerkin projection onto [net guess (while it pays), last %d "
                                                       "loop-1 solutions]" % RECYCLE}.get(head),
                   "option_overrides": args.opt or None},
        "forms": {k: {"iter_per_s": round(v["iter_per_s"], 3), "pressure_ms_per_step": round(v["pressure_ms_per_step"], 3),
                      "pressure_iters_per_step": v["pressure_iters_per_step"],
                      "learned_guesses_per_step": v["learned_guesses_per_step"],
                      "learned_guess_residual_ratio": v["learned_guess_residual_ratio"],
                      "deflation_starts_per_step": v["deflation_starts_per_step"],
                      "deflation_start_residual_ratio": v["deflation_start_residual_ratio"],
                      "deflation_vectors_per_start": v["deflation_vectors_per_start"],
                      "last_rel_residual": v["last_rel_residual"], "amg_setup": v["amg_setup"],
                      "stage_ms_per_step": v["stage_ms_per_step"]}
                  for k, v in forms.items()},
        "forms_skipped": skipped or None,
        "field_rel_error_vs_fp64": field_errors(forms),
        "device_memory_GB": {k: round(v / 1e9, 3) for k, v in h["memory"].items()},
        "mpnn_mfma": mfma_of(forms),
        "speedup_vs_fp64": (round(h["iter_per_s"] / forms["fp64"]["iter_per_s"], 3) if "fp64" in forms else None),
        "kernels": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in h["kernels"].items()},
        "refused_classes": h["refused_classes"],
        "roofline": roofline_of(h["roof_live"], args.m, probe),
        "cpu_baseline": cpu,
        "hbm_copy_probe": probe,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
