"""Batched cross-plane films: many independent instances in one launch per kernel.

The reference runs one MonteCarloSolver per process (src/main.cpp:86-102); a
parameter sweep over film thickness and background temperature is many such
processes.  Here every instance keeps its own device context (tables, carriers,
ledgers) and the contexts advance together through pmc_batch (include/pmc_capi.h):
each kernel of the step is launched once over all instances and the step is
replayed as a captured hipGraph.

The ensemble: thickness log-uniform in [10, 1000] nm along z, periodic 100 x 100
nm in x and y, 20 cells along z, background temperature uniform in [150, 450] K,
reservoirs at T_bg +/- 1 K, lifetimes at T_bg, initial linear profile, carriers
proportional to thickness (1e4 at 10 nm), the synthetic 16^3 x 6 Si-like table,
dt = 1 ps.
"""
import numpy as np

from phonomc_amd import runtime as rt

BASE = "c1_cross_plane.toml"


def film_ensemble(n, seed=7, carriers_at_10nm=10000, thickness_nm=(10.0, 1000.0), t_bg=(150.0, 450.0),
                  delta_t=1.0, steps=2000):
    """n instance descriptions: dicts with thickness (nm), t_bg (K), carriers, seed and the TOML text."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(thickness_nm[0]), np.log(thickness_nm[1])
    out = []
    for k in range(n):
        t = float(np.exp(rng.uniform(lo, hi)))
        tb = float(rng.uniform(*t_bg))
        n_car = int(round(carriers_at_10nm * t / 10.0))
        text = rt.config_text(BASE, **{
            "geometry.sizes": f"[100, 100, {t:.6f}]",
            "simulation.particle_count": n_car,
            "simulation.background_temperature": f"{tb:.6f}",
            "simulation.lifetime_temperature": f"{tb:.6f}",
            "simulation.random_seed": 1000 + k,
            "simulation.iterations": steps,
            "boundary.boundary_values": f"[{tb + delta_t:.6f}, {tb - delta_t:.6f}, 0, 0, 0, 0]",
        })
        out.append(dict(thickness_nm=t, t_bg=tb, carriers=n_car, seed=1000 + k, text=text))
    return out


def contexts(instances):
    """Problems and initialized device contexts of the instances."""
    problems = [rt.Problem(i["text"]) for i in instances]
    return problems, [rt.Context(p) for p in problems]


def kappa_eff(problem, ctx):
    """kappa_eff = -mean q_z / ((T_hot_side - T_cold_side) / L) with the reservoir
    temperatures at the film ends (MonteCarloSolver::update_heat_flux_and_conductivity,
    src/solver/Observables.cpp:38-142), from the context's current state (W/m/K)."""
    q = ctx.heat_flux()[:, 2]
    _, vol = problem.cells()
    f = problem.facets()
    res = [k for k, c in enumerate(f["conditions"]) if c == "T"]
    lo = min(res, key=lambda k: f["centroids"][k][2])
    hi = max(res, key=lambda k: f["centroids"][k][2])
    s, m, g, i, rough, rv = problem.views()
    length = (g.bmax[2] - g.bmin[2]) * 1e-10
    grad = (f["values"][hi] - f["values"][lo]) / length
    return -float((q * vol).sum() / vol.sum()) / grad


def throughput(n_instances=512, steps=50, warmup=5, seed=7, carriers_at_10nm=10000, sequential_steps=5, log=print):
    """Particle-steps/s of the ensemble stepped as one batch (graph replay and
    kernel-by-kernel) and context by context, all on the same instances."""
    import time

    inst = film_ensemble(n_instances, seed=seed, carriers_at_10nm=carriers_at_10nm)
    t0 = time.perf_counter()
    problems, ctx = contexts(inst)
    log(f"[batch] {n_instances} instances, {sum(i['carriers'] for i in inst):.4g} carriers, "
        f"built in {time.perf_counter() - t0:.1f} s")

    def carrier_steps():
        return sum(c.ledger()["carrier_steps"] for c in ctx)

    out = {"instances": n_instances, "carriers": int(sum(i["carriers"] for i in inst))}
    batch = rt.Batch(ctx, graph=True)
    batch.step(warmup)
    for name, graph, k in (("graph", True, steps), ("launches", False, steps)):
        lib = rt.lib()
        lib.pmc_batch_set_graph(batch.h, 1 if graph else 0)
        before = carrier_steps()
        t0 = time.perf_counter()
        batch.step(k)
        dt = time.perf_counter() - t0
        out[name] = {"value": (carrier_steps() - before) / dt, "ms_per_step": 1e3 * dt / k, "steps": k}
        log(f"[batch] {name}: {k} steps in {dt:.3f} s -> {out[name]['value']:.4e} particle-steps/s")
    batch.close()
    before = carrier_steps()
    t0 = time.perf_counter()
    for c in ctx:
        c.step(sequential_steps)
    ctx[-1].ledger()  # synchronizes
    dt = time.perf_counter() - t0
    out["one_by_one"] = {"value": (carrier_steps() - before) / dt, "ms_per_step": 1e3 * dt / sequential_steps,
                         "steps": sequential_steps}
    log(f"[batch] one context at a time: {sequential_steps} steps in {dt:.3f} s -> "
        f"{out['one_by_one']['value']:.4e} particle-steps/s")
    kap = np.array([kappa_eff(p, c) for p, c in zip(problems, ctx)])
    out["kappa_eff_range"] = [float(kap.min()), float(kap.max())]
    return out


def main():
    import argparse
    import json

    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--carriers-at-10nm", type=int, default=10000)
    args = ap.parse_args()
    print(json.dumps(throughput(args.instances, args.steps, args.warmup, carriers_at_10nm=args.carriers_at_10nm)))


if __name__ == "__main__":
    main()
