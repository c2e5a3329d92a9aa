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
"""
import argparse
import json
import os
import sys
import tempfile
import time

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
}
# Kernel names of each class in rocprofv3 output.
KERNEL_NAMES = {"transport": "k_transport", "sort": "k_sort", "moments": "k_moments", "relax": "k_relax"}
# Whole-step algorithmic bytes per particle-step: fused transport + tally pass
# (read and write pos 12, mode 4, occ 8, cell 4) and collision pass (read mode,
# cell, occ; write occ).
STEP_BYTES = 80


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(problem_text, steps):
    """Times the reference solver (oracle/_ref) on the full workload for `steps` steps."""
    from oracle import ref_ctypes
    from phonomc_amd import runtime as rt

    cores = host_cores()
    tmp = tempfile.mkdtemp(prefix="pmc_ref_")
    text = problem_text.replace('output_folder = "results/c2_inplane_film"', f'output_folder = "{tmp}/out"')
    path = os.path.join(tmp, "c2.toml")
    with open(path, "w") as f:
        f.write(text)
    problem = rt.Problem(text)
    if ref_ctypes.available():
        # The reference solver prints progress on stdout; keep stdout for the JSON line.
        sys.stdout.flush()
        saved = os.dup(1)
        os.dup2(2, 1)
        try:
            secs, ps = ref_ctypes.run(path, problem.material(), steps, threads=cores)
        finally:
            sys.stdout.flush()
            os.dup2(saved, 1)
            os.close(saved)
        kind = "reference"
    else:  # restatement when the reference build is absent
        from oracle import oracle_ctypes as O
        sim = O.Sim(problem, threads=cores)
        sim.initialize()
        t0 = time.perf_counter()
        before = sim.ledger()["carrier_steps"]
        sim.step(steps)
        secs = time.perf_counter() - t0
        ps = sim.ledger()["carrier_steps"] - before
        kind = "port"
    return {"value": ps / secs, "unit": "particle-steps/s", "cores": cores, "kind": kind,
            "sample": f"{WORKLOAD} at 1e7 carriers, {steps} steps of run_timestep after initialization, "
                      f"OpenMP threads={cores}", "seconds": secs}


def copy_probe(mib):
    """One device-to-device copy of `mib` MiB (hipMemcpy blit kernel): the
    calibration for HBM byte counters (FETCH_SIZE / WRITE_SIZE per dispatch)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--timing-steps", type=int, default=40)
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--no-cpu-baseline", action="store_true")
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

    from phonomc_amd import runtime as rt

    text = rt.config_text(WORKLOAD)
    problem = rt.Problem(text)
    ctx = rt.Context(problem)
    n0 = ctx.ledger()["live_carriers"]
    log(f"[bench] {WORKLOAD}: {n0} carriers initialized; warmup {args.warmup} steps")
    ctx.step(args.warmup)
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
        log(f"[bench] CPU baseline: reference solver, {args.cpu_steps} steps")
        baseline = cpu_baseline(text, args.cpu_steps)
        log(f"[bench] CPU baseline {baseline['value']:.4e} particle-steps/s on {baseline['cores']} threads")

    out = {
        "metric": "particle-steps/s @1 GPU",
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
        "config": {"workload": "in-plane film 1000x100 nm, 50x10 cells, rough 1 nm walls, 302/298 K reservoirs, "
                               "local lifetimes, 1e7 carriers, dt 0.5 ps",
                   "carriers": n0, "cells": int(problem.scalars()["cells"]), "time_step_ps": 0.5},
        "roofline": roofline,
        "cpu_baseline": baseline,
        "rta": {"evaluations": led1["newton_evals"] - led0["newton_evals"],
                "exact_carrier_sums": led1["newton_passes"] - led0["newton_passes"]},
    }
    if baseline:
        out["speedup_vs_cpu"] = value / baseline["value"]
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
