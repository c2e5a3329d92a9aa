#!/usr/bin/env python3
"""Throughput of the deviational Monte Carlo time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Workload: finite in-plane film, 1000 nm (x) by 100 nm (z), periodic in y,
50x10 cells, 1 nm rough top/bottom walls, reservoirs 302/298 K, local
lifetimes, 1e7 carriers, dt = 0.5 ps, synthetic Si-like 16^3 x 6 table.
A step is one call of the time step (transport, injection, temperature
refresh, RTA collisions, refresh, ledgers) over all carriers; the metric is
particle-steps/s = sum over timed steps of live carriers at step start / wall
time of the timed step loop, with all state resident in HBM.

Prints one JSON line.  The roofline object describes the kernel with the
largest share of step time (HIP-event durations on the step's stream); the
cpu_baseline object times the reference PhonoMC solver compiled from its own
sources (oracle/_ref) on a bounded sample of the same workload on the host
cores of this machine.

The warm_start object is the second half of the metric: the 3-D spectral
operator of the FinFET is trained on the device (BF16 MFMA) from the committed
data set of batched-solver fields (phonomc_amd/data/finfet_fields.npz), then
plain and warm-started Monte Carlo of a held-out case are timed, with common
random numbers, to 0.5 K RMS of a run started at the case's long-run field
(phonomc_amd/warmstart3d.py: evaluate).  Its roofline object is the
operator's GEMM rate (algorithmic flops over HIP-event durations of the GEMM
launches of forward + backward passes at the training batch) against the
dense BF16 MFMA peak.  --warm-start film runs the heated-film version instead.
"""
import argparse
import json
import math
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

# The CPUs this process may use, read before any OpenMP library loads: with a
# binding set, the OpenMP runtime pins the initial thread to one place and the
# affinity mask seen afterwards is that single place.
try:
    HOST_AT_START = set(os.sched_getaffinity(0))
except AttributeError:
    HOST_AT_START = None
# OpenMP reads its binding at start-up: threads of the CPU baseline bind one per
# physical core, packed (set before any OpenMP library loads).
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOAD = "c2_inplane_film.toml"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per carrier of each carrier-pass kernel (carrier arrays the
# pass must read and write in this design; tables and cell sums excluded).
BYTES_PER_CARRIER = {
    "transport": 16 + 16,  # read pos(12) mode(4); write pos(12) cell(4)
    "sort": 32 + 32,       # read pos, mode, occ, cell, uid (32); write them in cell order (32)
    "moments": 12 + 8,     # read mode(4) occ(8); write relaxation weight (8)
    "relax": 20 + 8,       # read mode(4) occ(8) weight(8); write occ(8)
    "callaway": 12 + 8,    # read mode(4) occ(8); write occ(8) (the Taylor terms stay in registers)
}
# Kernel names of each class in rocprofv3 output.
KERNEL_NAMES = {"transport": "k_transport", "sort": "k_sort", "moments": "k_moments", "relax": "k_relax",
                "callaway": "k_callaway"}
# Whole-step algorithmic bytes per particle-step: fused transport + tally pass
# (read and write pos 12, mode 4, occ 8, cell 4) and collision pass (read mode,
# cell, occ; write occ).
STEP_BYTES = 80
MFMA_BF16_PEAK_TFLOPS = 2500.0
WARM_DATA = os.path.join(ROOT, "phonomc_amd", "data", "heated_film_fields.npz")
WARM_TRAIN = 144  # the first 144 cases train, the rest are held out


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_topology():
    """Host CPUs this process may use: the affinity set, the physical cores among
    them (lscpu -p=CPU,Core,Socket) and the cgroup CPU quota when one is set.
    The baseline runs one OpenMP thread per physical core, capped by the quota."""
    if HOST_AT_START is not None:  # taken before OpenMP pinned this thread
        allowed = HOST_AT_START
    else:
        allowed = set(range(os.cpu_count() or 1))
    physical = None
    try:
        out = subprocess.run(["lscpu", "-p=CPU,Core,Socket"], capture_output=True, text=True, check=True).stdout
        cores = set()
        for line in out.splitlines():
            if line.startswith("#") or not line.strip():
                continue
            cpu, core, sock = line.split(",")[:3]
            if int(cpu) in allowed:
                cores.add((sock, core))
        physical = len(cores) or None
    except (OSError, subprocess.CalledProcessError, ValueError):
        pass
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    threads = physical or len(allowed)
    if quota:
        threads = min(threads, max(1, math.floor(quota)))
    return {"threads": max(1, threads), "physical_cores": physical, "affinity_cpus": len(allowed),
            "cgroup_cpus": quota, "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "omp_places": os.environ.get("OMP_PLACES")}


def cpu_baseline(problem_text, steps, workload=WORKLOAD, carriers=1e7, warmup=50):
    """Times the reference solver (oracle/_ref) on the full workload: `warmup`
    untimed steps after initialization, then `steps` timed steps."""
    from oracle import ref_ctypes
    from phonomc_amd import runtime as rt

    topo = cpu_topology()
    cores = topo["threads"]
    tmp = tempfile.mkdtemp(prefix="pmc_ref_")
    text = re.sub(r'output_folder = "[^"]*"', f'output_folder = "{tmp}/out"', problem_text)
    path = os.path.join(tmp, "workload.toml")
    with open(path, "w") as f:
        f.write(text)
    model = re.search(r'^model\s*=\s*"([^"]+)"', text, re.M)
    if model and model.group(1) != "box":  # the mesh file beside the input, where the reference looks for it
        shutil.copy(os.path.join(rt.CONFIGS, model.group(1)), tmp)
    problem = rt.Problem(text, input_directory=rt.CONFIGS)
    if ref_ctypes.available():
        # The reference solver prints progress on stdout; keep stdout for the JSON line.
        sys.stdout.flush()
        saved = os.dup(1)
        os.dup2(2, 1)
        try:
            lv = problem.layers()
            if lv is None:
                secs, ps = ref_ctypes.run(path, problem.material(), steps, threads=cores, warmup=warmup)
            else:  # layered materials: every layer's table, in order
                secs, ps = ref_ctypes.run_materials(path, [problem.layer_material(k) for k in range(lv.count)],
                                                    steps, threads=cores, warmup=warmup)[:2]
        finally:
            sys.stdout.flush()
            os.dup2(saved, 1)
            os.close(saved)
        kind = "reference"
    else:  # restatement when the reference build is absent
        from oracle import oracle_ctypes as O
        sim = O.Sim(problem, threads=cores)
        sim.initialize()
        sim.step(warmup)
        t0 = time.perf_counter()
        before = sim.ledger()["carrier_steps"]
        sim.step(steps)
        secs = time.perf_counter() - t0
        ps = sim.ledger()["carrier_steps"] - before
        kind = "port"
    return {"value": ps / secs, "unit": "particle-steps/s", "cores": cores, "kind": kind,
            "sample": f"{workload} at {carriers:.3g} carriers: {warmup} warm-up then {steps} timed steps of "
                      f"run_timestep after initialization, OpenMP threads={cores} (one per physical core, "
                      f"OMP_PROC_BIND={topo['omp_proc_bind']}, OMP_PLACES={topo['omp_places']})",
            "seconds": secs, "warmup_steps": warmup, "timed_steps": steps, "host": topo}


def copy_probe(mib):
    """One device-to-device copy of `mib` MiB (hipMemcpy blit kernel): the
    calibration for HBM byte counters (FETCH_SIZE / WRITE_SIZE per dispatch)."""
    import ctypes as C
    from phonomc_amd import runtime as rt
    # the HIP runtime the library is linked against (a bare "libamdhip64.so"
    # can name torch's bundled copy instead)
    hip = C.CDLL(os.path.join(rt.BUILD, "libpmc_device.so"))
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(n)) or hip.hipMalloc(C.byref(b), C.c_size_t(n)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(n))
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    rc = hip.hipMemcpy(b, a, C.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    dt = time.perf_counter() - t0
    hip.hipFree(a)
    hip.hipFree(b)
    if rc:
        raise RuntimeError(f"hipMemcpy failed: {rc}")
    log(f"[bench] copy probe: {mib} MiB in {dt * 1e3:.3f} ms ({2 * n / dt / 1e9:.1f} GB/s read+write)")


def pmc_summary(fetch_csv, write_csv, copy_mib, command):
    """HBM bytes per launch of each carrier-pass kernel from two rocprofv3
    --pmc runs (FETCH_SIZE, WRITE_SIZE in KiB per dispatch), each scaled by what
    the same counter reads for the device copy of copy_mib MiB in that run.
    Writes profiles/traffic_<kernel>.json."""
    import csv

    def per_kernel(path, counter):
        d = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            name = r["Kernel_Name"].split("(")[0].replace("pmc_dev::", "")
            d.setdefault(name, []).append(float(r["Counter_Value"]))
        return d

    fetch, write = per_kernel(fetch_csv, "FETCH_SIZE"), per_kernel(write_csv, "WRITE_SIZE")
    copy_bytes = float(copy_mib << 20)
    fetch_scale = copy_bytes / (max(fetch["__amd_rocclr_copyBuffer"]) * 1024.0)
    write_scale = copy_bytes / (max(write["__amd_rocclr_copyBuffer"]) * 1024.0)
    os.makedirs(os.path.join(ROOT, "profiles"), exist_ok=True)
    for cls, name in KERNEL_NAMES.items():
        if name not in fetch or name not in write:
            continue
        f = sum(fetch[name]) / len(fetch[name]) * 1024.0 * fetch_scale
        w = sum(write[name]) / len(write[name]) * 1024.0 * write_scale
        out = {"kernel": name, "dispatches": len(fetch[name]), "fetch_bytes": f, "write_bytes": w,
               "bytes_per_launch": f + w, "fetch_scale": fetch_scale, "write_scale": write_scale,
               "calibration": f"device copy of {copy_mib} MiB (__amd_rocclr_copyBuffer)", "command": command}
        with open(os.path.join(ROOT, "profiles", f"traffic_{cls}.json"), "w") as fh:
            json.dump(out, fh, indent=1)
        log(f"[bench] {name}: {f / 2**20:.1f} MiB read + {w / 2**20:.1f} MiB written per launch")


def gemm_rate(model, x, y, reps=50):
    """Roofline object of the operator's GEMMs: forward + backward passes at the
    training batch (8 cases), algorithmic flops over HIP-event GEMM durations."""
    from phonomc_amd import warmstart as ws

    op = model.op
    op.set_timing(True)
    op.gemm_stats(reset=True)
    for _ in range(reps):
        op.loss_grad(x[:8], y[:8] / ws.Y_SCALE)
    flops, launches, gemm_ms = op.gemm_stats(reset=True)
    op.set_timing(False)
    tflops = flops / (gemm_ms / 1e3) / 1e12
    return {"bound": "mfma", "achieved": tflops, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": tflops / MFMA_BF16_PEAK_TFLOPS, "traffic": None, "kernel": "k_gemm",
            "flops_per_pass": flops / reps, "launches_per_pass": launches / reps,
            "gemm_ms_per_pass": gemm_ms / reps, "avg_launch_ms": gemm_ms / max(1, launches)}


def warm_start_leg(epochs=None, evaluate=True):
    import numpy as np

    from phonomc_amd import warmstart as ws

    params, x, y = ws.load_dataset(WARM_DATA)
    model = ws.WarmStart()
    t0 = time.perf_counter()
    losses = model.fit(x[:WARM_TRAIN], y[:WARM_TRAIN], **({"epochs": epochs} if epochs else {}))
    train_s = time.perf_counter() - t0
    pred = model.predict(x[WARM_TRAIN:])
    held = np.sqrt(np.mean((pred - y[WARM_TRAIN:]) ** 2, axis=1))
    without = np.sqrt(np.mean(y[WARM_TRAIN:] ** 2, axis=1))
    log(f"[bench] operator trained in {train_s:.1f} s: held-out RMS {held.mean():.3f} K "
        f"(reservoir field {without.mean():.3f} K)")
    roofline = gemm_rate(model, x, y)
    log(f"[bench] operator GEMMs: {roofline['achieved']:.3f} TFLOP/s, {roofline['launches_per_pass']:.0f} launches "
        f"per forward + backward, {roofline['avg_launch_ms'] * 1e3:.1f} us per launch")
    if not evaluate:
        return {"held_out_rms_k": float(held.mean()), "train_seconds": train_s, "train_epochs": int(len(losses)),
                "roofline": roofline}
    k = int(np.argmax(y[WARM_TRAIN:].max(axis=1)))
    case = ws.cases_from_params(params[WARM_TRAIN:])[k]
    ev = ws.evaluate(model, case, particles=1000000, log=log)
    return {
        "workload": "heated film 500x100 nm, 50x10 cells, Gaussian line source, reservoirs 295-305 K, "
                    "1e6 carriers, dt 0.5 ps",
        "tol_k": ev["tol_k"], "window_steps": ev["window_steps"], "particles": ev["particles"],
        "case": ev["case"], "peak_rise_k": ev["peak_rise_k"],
        "plain_steps_to_steady": None if ev["plain_windows"] is None else ev["plain_windows"] * ev["window_steps"],
        "warm_steps_to_steady": None if ev["warm_windows"] is None else ev["warm_windows"] * ev["window_steps"],
        "plain_seconds": ev["plain_seconds"], "warm_seconds": ev["warm_seconds"],
        "speedup_steps": ev.get("speedup_steps"), "speedup_wall": ev.get("speedup_wall"),
        "operator_rms_k": ev["surrogate_rms_k"], "uniform_start_rms_k": ev["uniform_rms_k"],
        "plain_late_rms_k": ev["plain_late_rms_k"], "warm_late_rms_k": ev["warm_late_rms_k"],
        "held_out_rms_k": float(held.mean()), "held_out_cases": int(len(held)),
        "train_cases": WARM_TRAIN, "train_seconds": train_s, "train_epochs": int(len(losses)),
        "roofline": roofline,
    }


FINFET_EVAL = dict(carriers=2000000, ref_burn_in=3000, ref_average=2000, steps=3000, window=20, sample=5)


def finfet_warm_start_leg(epochs=48, evaluate=True):
    """The FinFET warm start: the 3-D operator (64 x 32 x 32 grid, 3 channels, 4
    spectral layers, width 32, 12 modes per axis, BF16 MFMA) trained on the
    device from the committed data set of batched-solver fields
    (phonomc_amd/data/finfet_fields.npz), then one held-out case (the median
    rise of the held-out set) run plain and warm-started with common random
    numbers to 0.5 K RMS of a run started at its long-run field
    (phonomc_amd/warmstart3d.py: evaluate)."""
    import numpy as np

    from phonomc_amd import warmstart3d as w3

    model, info, (x, y, w) = w3.train(epochs=epochs, log=log)
    roofline = w3.gemm_roofline(model, x, y, w)
    log(f"[bench] operator GEMMs: {roofline['achieved']:.2f} TFLOP/s over {roofline['launches_per_pass']:.0f} launches "
        f"per forward + backward ({roofline['avg_launch_ms'] * 1e3:.1f} us per launch)")
    out = {"workload": "FinFET (fin 10x40x100 nm on a 200x100x50 nm substrate, 64x32x32 grid, 23 536 cells), "
                       "linearized Callaway, Gaussian source, bottom reservoir at t_bg, dt 1 ps",
           "operator": {"grid": [64, 32, 32], "channels": 3, "layers": 4, "width": 32, "modes": 12, **info},
           "roofline": roofline}
    if not evaluate:
        return out
    cases = w3.load_dataset_cases(w3.DATA_PATH)[-w3.HELD_OUT:]
    rms = np.sqrt((y[-w3.HELD_OUT:].astype(np.float64) ** 2).sum(axis=1))
    case = cases[int(np.argsort(rms)[len(rms) // 2])]
    e = FINFET_EVAL
    t0 = time.perf_counter()
    t_ref = w3.reference_fields([case], e["carriers"], e["ref_burn_in"], e["ref_average"], log=log)[0]
    log(f"[bench] FinFET T_ref in {time.perf_counter() - t0:.1f} s")
    ev = w3.evaluate(model, case, e["carriers"], t_ref, e["steps"], e["window"], e["sample"], log=log)
    out.update({k: ev[k] for k in ("case", "carriers", "window_steps", "tol_k", "steady_rms_rise_k", "peak_rise_k",
                                   "operator_rms_k", "operator_rel_rms", "plain_steps_to_steady",
                                   "warm_steps_to_steady", "plain_seconds", "warm_seconds", "predict_seconds")})
    out["speedup_steps"] = ev.get("speedup_steps")
    out["speedup_wall"] = ev.get("speedup_wall")
    out["reference"] = {"burn_in_steps": e["ref_burn_in"], "average_steps": e["ref_average"]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--timing-steps", type=int, default=40)
    ap.add_argument("--cpu-steps", type=int, default=200, help="timed steps of the CPU baseline")
    ap.add_argument("--cpu-warmup", type=int, default=50, help="untimed steps of the CPU baseline first")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-warm-start", action="store_true", help="skip the warm-start leg")
    ap.add_argument("--warm-start", choices=["finfet", "film"], default="finfet",
                    help="warm-start leg: the FinFET operator (default) or the heated film")
    ap.add_argument("--workload", default=WORKLOAD, help="bundled TOML input (phonomc_amd/configs) to time")
    ap.add_argument("--particles", type=int, default=0, help="override simulation.particle_count")
    ap.add_argument("--autotune", type=int, default=0, metavar="STEPS",
                    help="after warm-up, run pmc_autotune with STEPS steps per candidate and keep its choice")
    ap.add_argument("--surrogate-only", type=int, metavar="EPOCHS", default=0,
                    help="only train the operator for EPOCHS epochs and print its GEMM roofline (for profiling)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run one device copy of this size first (byte-counter calibration)")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write profiles/traffic_*.json from rocprofv3 counter CSVs and exit")
    ap.add_argument("--pmc-command", default="", help="command recorded with --pmc-summary")
    args = ap.parse_args()
    if args.pmc_summary:
        pmc_summary(args.pmc_summary[0], args.pmc_summary[1], args.copy_probe_mib or 1024, args.pmc_command)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (multi-GPU sharding is "
              f"a design note, not part of this build).", file=sys.stderr)
        return 2

    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
    if args.surrogate_only:
        leg = warm_start_leg if args.warm_start == "film" else finfet_warm_start_leg
        print(json.dumps(leg(epochs=args.surrogate_only, evaluate=False)), flush=True)
        return 0

    from phonomc_amd import runtime as rt

    text = rt.config_text(args.workload, **({"simulation.particle_count": args.particles} if args.particles else {}))
    problem = rt.Problem(text, input_directory=rt.CONFIGS)
    ctx = rt.Context(problem)
    n0 = ctx.ledger()["live_carriers"]
    log(f"[bench] {args.workload}: {n0} carriers initialized; warmup {args.warmup} steps")
    ctx.step(args.warmup)
    tuned = None
    if args.autotune:
        chosen, report = ctx.autotune(args.autotune)
        tuned = {"chosen": chosen, "kernel_ms": report}
        log(f"[bench] autotune: {chosen} {report}")
    led0 = ctx.ledger()  # synchronizes the device
    t0 = time.perf_counter()
    ctx.step(args.steps)  # returns after the device finished the last step
    t1 = time.perf_counter()
    led1 = ctx.ledger()
    elapsed = t1 - t0
    carrier_steps = led1["carrier_steps"] - led0["carrier_steps"]
    value = carrier_steps / elapsed
    log(f"[bench] {args.steps} steps in {elapsed:.4f} s -> {value:.4e} particle-steps/s")

    # Per-kernel durations with HIP events on the context stream.
    ctx.set_timing(True)
    ctx.reset_kernel_times()
    ctx.step(args.timing_steps)
    times = ctx.kernel_times()
    ctx.set_timing(False)
    live = ctx.ledger()["live_carriers"]
    candidates = {k: v for k, v in times.items() if k in BYTES_PER_CARRIER and v[1] > 0}
    dominant = max(candidates, key=lambda k: candidates[k][0])
    total_ms, launches = times[dominant]
    avg_s = total_ms / launches / 1e3
    traffic_bytes = BYTES_PER_CARRIER[dominant] * live
    achieved = traffic_bytes / avg_s / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dominant}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")
    step_gbs = value * STEP_BYTES / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dominant,
                "bytes_per_carrier": BYTES_PER_CARRIER[dominant], "carriers_per_launch": live,
                "avg_launch_ms": total_ms / launches,
                "step": {"bytes_per_particle_step": STEP_BYTES, "achieved": step_gbs,
                         "frac": step_gbs / HBM_PEAK_GBS},
                "step_breakdown_ms": {k: v[0] / max(1, times["step"][1]) for k, v in times.items() if v[1]}}
    log(f"[bench] kernel times (ms/step): {json.dumps(roofline['step_breakdown_ms'])}")

    baseline = None
    if not args.no_cpu_baseline:
        log(f"[bench] CPU baseline: reference solver, {args.cpu_warmup} + {args.cpu_steps} steps")
        baseline = cpu_baseline(text, args.cpu_steps, args.workload, n0, warmup=args.cpu_warmup)
        log(f"[bench] CPU baseline {baseline['value']:.4e} particle-steps/s on {baseline['cores']} threads")

    ctx.close()
    warm = None
    if not args.no_warm_start:
        log("[bench] warm start: training the operator and timing plain vs warm-started runs")
        from phonomc_amd import warmstart3d as w3
        film = args.warm_start == "film" or not os.path.exists(w3.DATA_PATH)
        if film and args.warm_start != "film":
            log(f"[bench] {w3.DATA_PATH} absent: the heated-film warm start runs instead")
        warm = warm_start_leg() if film else finfet_warm_start_leg()
        log(f"[bench] warm start: {warm.get('plain_steps_to_steady')} vs {warm.get('warm_steps_to_steady')} steps to "
            f"0.5 K RMS ({warm.get('speedup_wall')} x wall); GEMM {warm['roofline']['achieved']:.2f} TFLOP/s")

    out = {
        "metric": "particle-steps/s @1 GPU; time-to-steady T (0.5 K RMS) warm-start vs plain",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 flight / f64 tallies",
        "data": "synthetic (seeded Si-like 16^3x6 phonon table, Philox-sampled carriers)",
        "config": {"workload": ("in-plane film 1000x100 nm, 50x10 cells, rough 1 nm walls, 302/298 K reservoirs, "
                                "local lifetimes, 1e7 carriers, dt 0.5 ps") if args.workload == WORKLOAD and
                               not args.particles else f"{args.workload} ({n0} carriers)",
                   "carriers": n0, "cells": int(problem.scalars()["cells"]),
                   "time_step_ps": problem.views()[0].time_step_ps},
        "roofline": roofline,
        "cpu_baseline": baseline,
        "rta": {"evaluations": led1["newton_evals"] - led0["newton_evals"],
                "exact_carrier_sums": led1["newton_passes"] - led0["newton_passes"]},
    }
    if baseline:
        out["speedup_vs_cpu"] = value / baseline["value"]
    out["warm_start"] = warm
    if tuned is not None:
        out["autotune"] = tuned
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
