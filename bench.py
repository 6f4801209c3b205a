"""Throughput of the MI355X SPH impact step (one JSON line on stdout).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload layered_4m]
                    [--preroll P] [--run-steps T] [--no-cpu] [--no-fp64] [--no-ensemble]

A step is one System::step of every particle of the workload (kick/drift, cell
sort, adaptive-h density, EOS, strain + forces, damage, finish), with the CFL
time step kept on the device.  The default workload is the ~4.0M-particle
layered impact (iron sphere R = 60 dx at 2 deg and 8 km/s into 192x96 columns
of 32 water cells over 136 granite cells, strength and Weibull damage on, 1000
steps), mixed precision.  Inputs are resident in HBM before any timed region.

Schedule of each precision's run: init, P steps that carry the impact well past
contact (default: half the workload's step count), the autotuner on that
post-contact state, W warm-up steps, then the K timed steps (replayed from a
captured hipGraph and bracketed by device synchronisation), a few profiled
steps, and the rest of the workload's steps.  `value` is the K-step window;
`run` reports the whole workload (every step but the profiled ones timed).

Besides the contract fields the line carries:
  roofline     -- dominant kernel's algorithmic HBM bytes / its mean HIP-event
                  time over the profiled steps, its measured traffic and pipe
                  utilisation from the committed rocprof summaries of these
                  device sources, and its algorithmic FP32 rate;
  cpu_baseline -- the reference solver's own code (oracle/_ref, compiled from the
                  reference headers) timed on this host: init and at least 3 steps
                  of the whole workload;
  fp64         -- the all-FP64 device path on the same schedule and its ratio;
  ensemble_256 -- runs/hr and step-level HBM fraction of the 256-impact batch.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from deorbit_amd import dsph, ics, roofline  # noqa: E402

WORKLOADS = {"impact_37k": ics.impact_37k, "graze_1m": ics.graze_1m, "layered_4m": ics.layered_4m,
             "ensemble_256": None, "halfspace_32m": ics.halfspace_32m}
WORKLOAD = {
    "impact_37k": "impact_37k: R=10dx iron sphere at 45 deg, 8 km/s, onto a 64x32x16 granite slab, no strength",
    "graze_1m": "graze_1m: R=40dx iron sphere at 3 deg, 8 km/s, into a 128x64x96 granite slab, strength+damage",
    "layered_4m": "layered_4m: R=60dx iron sphere at 2 deg, 8 km/s, into 192x96 columns of 32 water over 136 "
                  "granite",
    "ensemble_256": "ensemble_256: 256 independent iron-on-granite impacts in one batch (N logU[8k,200k], "
                    "1-45 deg, 4-12 km/s, impactor/target 0.05-0.3, strength+damage), 500 steps per instance",
    "halfspace_32m": "halfspace_32m: R=120dx iron sphere at 1.5 deg, 8 km/s, onto a 384x192x339 granite "
                     "half-space",
}
ENSEMBLE_STEPS = 500       # steps per instance of the batched ensemble


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_run(sysm, steps, chunk=50):
    """Wall time of `steps` steps, enqueued `chunk` steps per dsph_run call (each call
    waits for its steps and returns the device flags), so a long run never queues
    more than a few thousand kernels ahead of the device."""
    if steps <= 0:
        return 0.0
    sysm.synchronize()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        sysm.run(k, 0.0, GRAPH)
        done += k
    sysm.synchronize()
    return time.perf_counter() - t0


def profile_leg(s, steps):
    s.set_profiling(True)
    s.reset_profile()
    s.run(steps, 0.0, False)
    s.synchronize()
    prof = s.profile()
    s.set_profiling(False)
    return prof


GRAPH = True     # steps replayed from a captured hipGraph (--no-graph: stream launches)


def gpu_leg(ic, precision, preroll, warmup, steps, run_steps, autotune=True, profile_steps=0, tuning=None):
    """One precision's run on the schedule of the module docstring.  Returns the
    system, the K-step time, the profile of the profiled steps and the whole-run
    timing.  `tuning` (a dict of dsph_tuning fields) replaces the autotuner."""
    name = "fp64" if precision == dsph.FP64 else "mixed"
    s = dsph.make_batch(ic, precision) if isinstance(ic, list) else dsph.make_system(ic, precision)
    s.init()
    t_pre = timed_run(s, preroll)
    d = s.diagnostics()
    log(f"{name}: {preroll} steps before the window ({t_pre:.2f} s), t={d['t'] * 1e3:.2f} ms, "
        f"umax={d['umax']:.3g} J/kg, dmax={d['dmax']:.3f}")
    if tuning:
        s.set_tuning(**tuning)
        log(f"tuning ({name}, at step {preroll}): {s.tuning()}")
    elif autotune:
        chosen = s.autotune(3)
        log(f"autotune ({name}, at step {preroll}): {chosen}")
    t_warm = timed_run(s, warmup)
    el = timed_run(s, steps)
    prof = profile_leg(s, profile_steps) if profile_steps else None
    rest = max(0, run_steps - preroll - warmup - steps - profile_steps)
    t_rest = timed_run(s, rest)
    timed = preroll + warmup + steps + rest
    run = {"steps": preroll + warmup + steps + profile_steps + rest, "timed_steps": timed,
           "ms_per_step": 1e3 * (t_pre + t_warm + el + t_rest) / max(1, timed),
           "t_sim_end": s.diagnostics()["t"]}
    return s, el, prof, run


def host_cores() -> dict:
    """CPU model, logical CPUs, the CPUs this process may run on and its cgroup
    CPU quota: the host threads a CPU solver can actually use here."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            if q != "max":
                quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    if quota is not None:
        usable = max(1, min(usable, int(quota)))
    return {"cpu_model": model, "nproc": nproc, "usable_cores": usable, "cgroup_cpu_quota": quota}


def time_cpu_solver(ic, which, nthreads, seconds_budget, min_steps=3, max_steps=20):
    """Init, then as many steps as fit the budget (at least min_steps); median step time."""
    from oracle import pyoracle as po
    o = po.make_system(ic, ic["params"], ic["materials"], which, nthreads=nthreads,
                       seed_damage=12345 if ic["params"]["damage_on"] else None, dx=ic["dx"])
    t0 = time.perf_counter()
    o.init()
    t_init = time.perf_counter() - t0
    times = []
    while True:
        dt = o.timestep()
        t0 = time.perf_counter()
        o.step(dt)
        times.append(time.perf_counter() - t0)
        if len(times) >= max_steps or (len(times) >= min_steps and sum(times) + t_init > seconds_budget):
            break
    return float(np.median(times)), len(times), t_init


def cpu_baseline(ic, seconds_budget=30.0):
    """The reference solver on the host with its own thread cap (min(16, hardware
    threads), sph.hpp:58-60); when this process may use more cores than that, the
    oracle restatement (bit-identical to the reference) with the cap lifted too."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import pyoracle as po
    host = host_cores()
    kind = "reference" if os.path.exists(po.REF_SO) else "port"
    which = "ref" if kind == "reference" else "oracle"
    cores = min(16, host["nproc"])
    med, nsteps, t_init = time_cpu_solver(ic, which, cores, seconds_budget)
    out = {"value": len(ic["mass"]) / med, "unit": "particle-steps/s", "cores": cores, "kind": kind,
           "sample": f"init and {nsteps} step(s) of the whole workload ({len(ic['mass'])} particles), median "
                     f"{med * 1e3:.0f} ms/step (init {t_init:.1f} s excluded); std::thread parallel_for over "
                     f"{cores} host threads",
           "host": host}
    if host["usable_cores"] > cores:
        med2, n2, _ = time_cpu_solver(ic, "oracle", host["usable_cores"], seconds_budget)
        out["all_cores"] = {"value": len(ic["mass"]) / med2, "cores": host["usable_cores"], "kind": "port",
                            "sample": f"{n2} step(s), median {med2 * 1e3:.0f} ms/step"}
    else:
        out["all_cores"] = {"value": out["value"], "cores": cores, "kind": kind,
                            "sample": f"this process may use {host['usable_cores']} cores, no more than the "
                                      "reference's own cap: the capped figure is the all-cores figure"}
    return out


def ensemble_leg(warmup, steps, preroll, autotune=True):
    """The 256-impact batch: runs/hr from the K-step window after `preroll` steps,
    and the step-level algorithmic HBM fraction."""
    t0 = time.perf_counter()
    ic = ics.ensemble(256, ENSEMBLE_STEPS)
    n = sum(len(m["mass"]) for m in ic)
    log(f"ensemble of {len(ic)} impacts: n={n} (built in {time.perf_counter() - t0:.1f} s)")
    s, el, _, _ = gpu_leg(ic, dsph.MIXED, preroll, warmup, steps, 0, autotune)
    s.close()
    value = n * steps / el
    return {"runs_per_hour": len(ic) * 3600.0 / (ENSEMBLE_STEPS * el / steps), "particles": n,
            "particle_steps_per_s": value, "ms_per_step": 1e3 * el / steps, "steps": steps,
            "window": f"steps {preroll + warmup + 1}-{preroll + warmup + steps} of {ENSEMBLE_STEPS}",
            "step_hbm_frac": round(roofline.step_bytes("mixed") * value / 1e9 / roofline.MI355X_HBM_GBPS, 5)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="layered_4m", choices=list(WORKLOADS))
    ap.add_argument("--preroll", type=int, default=None,
                    help="steps before the autotuner and the timed window (default: half the workload's steps)")
    ap.add_argument("--run-steps", type=int, default=None,
                    help="steps of the whole run (default: the workload's step count; 0 = stop after the window)")
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-fast", action="store_true", help="skip the DSPH_MIXED_FAST leg")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every step's kernels on the stream instead of replaying a hipGraph "
                         "(for tracing tools)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time the library's default tuning instead of the autotuner's choice")
    ap.add_argument("--tuning", default=None,
                    help="fixed tuning instead of the autotuner, e.g. cell_factor=1.45,block_pair=64 (profiling runs: "
                         "counter passes then time the same configuration)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy kernel (counter calibration)")
    args = ap.parse_args()
    global GRAPH
    GRAPH = not args.no_graph
    if args.copy_probe:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": a.numel() * 4, "copy_probe_bytes_written": b.numel() * 4,
                          "launches": 3}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: the SPH step has no multi-GPU path in this build "
              "(see the sharding design); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "ensemble_256":
        ic = ics.ensemble(256, ENSEMBLE_STEPS)
        n = sum(len(m["mass"]) for m in ic)
        cfg_steps = ENSEMBLE_STEPS
        log(f"workload ensemble of {len(ic)} impacts: n={n}")
        cpu_ic = sorted(ic, key=lambda m: len(m["mass"]))[len(ic) // 2]     # the median instance
    else:
        ic = WORKLOADS[args.workload]()
        n = len(ic["mass"])
        cfg_steps = int(ic["meta"]["steps"])
        log(f"workload {ic['meta']['name']}: n={n}, {cfg_steps} steps")
        cpu_ic = ic
    preroll = cfg_steps // 2 if args.preroll is None else args.preroll
    run_steps = cfg_steps if args.run_steps is None else args.run_steps
    tune = not args.no_autotune
    fixed = None
    if args.tuning:
        fixed = {}
        for kv in args.tuning.split(","):
            k, v = kv.split("=")
            fixed[k.strip()] = float(v) if k.strip() == "cell_factor" else int(v)

    s, el, prof, run = gpu_leg(ic, dsph.MIXED, preroll, args.warmup, args.steps, run_steps, tune,
                               args.profile_steps, fixed)
    value = n * args.steps / el
    ms = 1e3 * el / args.steps
    log(f"mixed: {ms:.3f} ms/step  {value / 1e6:.1f} M particle-steps/s (steps {preroll + args.warmup + 1}-"
        f"{preroll + args.warmup + args.steps}); whole run {run['ms_per_step']:.3f} ms/step over "
        f"{run['timed_steps']} steps")
    for k, (tot, cnt) in sorted(prof.items(), key=lambda kv: -kv[1][0]):
        if cnt:
            log(f"  {k:14s} {tot / args.profile_steps:8.3f} ms/step  ({cnt} launches)")
    # the two pair kernels against their rooflines; the dominant one (time per step)
    # is the line's roofline.  Algorithmic bytes and flops per particle and step
    # (deorbit_amd/roofline.py); the density kernel runs once per possible round
    # (4 launches per step, usually one active), so its figures are per step:
    # bytes per step / time per step = bytes per launch / mean launch time.
    pdir = os.path.join(ROOT, "profiles")
    ps = args.profile_steps
    # the force pass over neighbour lists or by cell walks (whichever the tuning
    # selected); the benchmark workloads have no periodic axis: non-periodic variants,
    # instantiated for one segment except on the batched ensemble
    fphase = "force_list" if prof.get("force_list", (0.0, 0))[0] > prof["force"][0] else "force"
    one = "false" if args.workload == "ensemble_256" else "true"
    kernels = {
        "force": dict(kernel="k_force_list<float, false>" if fphase == "force_list" else f"k_force<float, false, {one}>",
                      phase=fphase, bytes_pp=roofline.kernel_bytes("mixed")["force"], flops_pp=roofline.force_flops()),
        "density": dict(kernel=f"k_density_round_hp<false, {one}>", phase="density",
                        bytes_pp=roofline.density_round_bytes("mixed", True), flops_pp=roofline.density_flops()),
    }
    for key, k in kernels.items():
        tot, launches = prof[k["phase"]]
        step_ms = tot / ps
        achieved = k["bytes_pp"] * n / (step_ms * 1e-3) / 1e9
        tflops = k["flops_pp"] * n / (step_ms * 1e-3) / 1e12
        # HBM bytes per launch and pipe utilisation from the committed rocprof summaries
        # of this kernel and workload, used only when recorded for these device sources
        rec = roofline.recorded_traffic(pdir, k["kernel"], args.workload)
        pipes = roofline.recorded_pipes(pdir, k["kernel"], args.workload)
        lps = launches / ps
        k.update(step_ms=round(step_ms, 4), launches_per_step=lps, mean_launch_ms=round(tot / max(1, launches), 4),
                 achieved=round(achieved, 2), frac=round(achieved / roofline.MI355X_HBM_GBPS, 5),
                 traffic=None if rec is None else rec["bytes"], pipes=pipes, bound=roofline.limiter(pipes),
                 valu_tflops=round(tflops, 3), valu_frac=round(tflops / roofline.MI355X_FP32_TFLOPS, 5))
    dom = max(kernels, key=lambda key: kernels[key]["step_ms"])
    kd = kernels[dom]
    # per launch: algorithmic bytes (bytes per step / launches per step) over the mean launch
    # time give the same rate as per step; traffic is the measured HBM bytes per launch
    roof = {"bound": kd["bound"], "kernel": kd["kernel"], "achieved": kd["achieved"],
            "peak": roofline.MI355X_HBM_GBPS, "unit": "GB/s", "frac": kd["frac"], "traffic": kd["traffic"],
            "algorithmic_bytes": kd["bytes_pp"] * n / kd["launches_per_step"],
            "bytes_per_particle_step": kd["bytes_pp"],
            "launches_per_step": kd["launches_per_step"], "mean_launch_ms": kd["mean_launch_ms"],
            "source_digest": roofline.source_digest(), "pipes": kd["pipes"],
            "valu": {"flops_per_particle_step": kd["flops_pp"], "achieved_tflops": kd["valu_tflops"],
                     "peak_tflops": roofline.MI355X_FP32_TFLOPS, "frac": kd["valu_frac"]},
            "step_frac": round(roofline.step_bytes("mixed") * value / 1e9 / roofline.MI355X_HBM_GBPS, 5),
            "step_valu_frac": round(roofline.step_flops() * value / 1e12 / roofline.MI355X_FP32_TFLOPS, 5),
            "kernels": {key: {f: k[f] for f in ("kernel", "step_ms", "launches_per_step", "mean_launch_ms",
                                                "bytes_pp", "achieved", "frac", "traffic", "bound", "pipes",
                                                "valu_tflops", "valu_frac")} for key, k in kernels.items()}}
    log(f"roofline: dominant {kd['kernel']} {kd['step_ms']:.3f} ms/step, {kd['achieved']:.0f} GB/s algorithmic "
        f"({100 * kd['frac']:.2f} % of HBM), {kd['valu_tflops']:.1f} TFLOP/s; bound {kd['bound']}")
    diag = s.diagnostics()
    s.close()

    fp64 = None
    if not args.no_fp64:
        s64, el64, _, run64 = gpu_leg(ic, dsph.FP64, preroll, args.warmup, args.steps, run_steps, tune)
        v64 = n * args.steps / el64
        s64.close()
        fp64 = {"value": v64, "ms_per_step": 1e3 * el64 / args.steps, "mixed_over_fp64": value / v64,
                "run_ms_per_step": run64["ms_per_step"],
                "run_mixed_over_fp64": run64["ms_per_step"] / run["ms_per_step"] if run["timed_steps"] else None}
        log(f"fp64: {fp64['ms_per_step']:.3f} ms/step  ratio {fp64['mixed_over_fp64']:.2f}x; whole run "
            f"{run64['ms_per_step']:.3f} ms/step")

    fast = None
    if not args.no_fast:
        sf, elf, _, runf = gpu_leg(ic, dsph.MIXED_FAST, preroll, args.warmup, args.steps, run_steps, tune)
        sf.close()
        vf = n * args.steps / elf
        fast = {"value": vf, "ms_per_step": 1e3 * elf / args.steps, "run_ms_per_step": runf["ms_per_step"],
                "precision": "FP32 density sums and pressure as well (per-particle P ~5e-3 near rho0)"}
        if fp64 is not None:
            fast["over_fp64"] = vf / fp64["value"]
        log(f"mixed_fast: {fast['ms_per_step']:.3f} ms/step; whole run {runf['ms_per_step']:.3f} ms/step")

    ens = None
    if not args.no_ensemble and args.workload != "ensemble_256":
        ens = ensemble_leg(args.warmup, args.steps, ENSEMBLE_STEPS // 2, tune)
        log(f"ensemble_256: {ens['ms_per_step']:.3f} ms/step, {ens['runs_per_hour']:.0f} runs/hr, "
            f"step HBM fraction {ens['step_hbm_frac']:.4f}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cpu_ic)
        if cpu_ic is not ic:
            cpu["sample"] = f"median-size ensemble instance ({len(cpu_ic['mass'])} particles): " + cpu["sample"]
        log(f"cpu: {cpu['value'] / 1e6:.3f} M particle-steps/s ({cpu['kind']}, {cpu['cores']} threads); "
            f"all {cpu['all_cores']['cores']} usable cores: {cpu['all_cores']['value'] / 1e6:.3f} M/s; "
            f"{cpu['host']['cpu_model']}, nproc {cpu['host']['nproc']}")

    line = {
        "metric": "SPH particle-steps/s on 1 MI355X, mixed vs FP64 at stated error; runs/hr",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 force pass, stress, sound speed; f64 density terms, pressure, state, reductions",
        "data": "synthetic (seeded lattice impact IC)",
        "config": {"workload": WORKLOAD[args.workload], "particles": n, "dt": "device CFL (cfl=0.2)",
                   "graph": GRAPH, "window": f"steps {preroll + args.warmup + 1}-{preroll + args.warmup + args.steps}"
                   f" of {cfg_steps} (tuning {args.tuning} from step {preroll})" if fixed else
                   f"steps {preroll + args.warmup + 1}-{preroll + args.warmup + args.steps}"
                   f" of {cfg_steps} (autotuned at step {preroll})" if tune else
                   f"steps {preroll + args.warmup + 1}-{preroll + args.warmup + args.steps} of {cfg_steps}",
                   "t_sim_end": diag["t"]},
        "run": run, "roofline": roof, "cpu_baseline": cpu, "fp64": fp64, "mixed_fast": fast,
    }
    if ens is not None:
        line["ensemble_256"] = ens
    if args.workload == "ensemble_256":
        # every batched step advances all 256 instances by one step
        line["runs_per_hour"] = len(ic) * 3600.0 / (ENSEMBLE_STEPS * el / args.steps)
        line["config"]["instances"] = len(ic)
        line["config"]["steps_per_instance"] = ENSEMBLE_STEPS
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
