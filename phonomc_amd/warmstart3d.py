"""Learned warm start of the FinFET (the fin on its substrate of
configs/finfet_callaway.toml): a three-dimensional spectral operator
(phonomc_amd.fno, BF16 matrix cores) maps the heat-source density, the solid
mask and the substrate (reservoir) temperature on the 64 x 32 x 32 voxel grid
to the steady temperature rise; its prediction seeds the initial carrier
population through the per-cell initial field (pmc_init.temperature_mode = 2,
the role src/solver/Initialization.cpp:181-265 gives the uniform and linear
fields).  Plain Monte Carlo from the background temperature stays the fallback
and the yardstick.

A case (sample_cases) is
    height   source centre above the fin base (nm, 0-100; the fin is 100 nm tall)
    sigma    isotropic Gaussian width (nm, 3-15)
    power    peak power density (W/m^3, 1e19-1e21, log-uniform)
    t_bg     substrate temperature = background = bottom reservoir (K, 250-400)
on linearized Callaway transport with the config-3 physics (N/U + isotope,
diffuse walls), dt = 1 ps.  Linearized transport is linear in the source: a
case's rise is its total source power times a field that depends only on the
source shape, position and t_bg.  The operator therefore sees the source
density divided by its peak and predicts the rise per microwatt of source
power; the prediction is scaled by the case's power.

Data come from the batched runner (pmc_batch): every case is an instance of one
batch stepped `burn_in` steps from t_bg, then averaged over `average` steps.
Time to steady is measured as for the heated film (phonomc_amd.warmstart):
window averages against a long plain run's field, 0.5 K RMS over the cells.
"""
import concurrent.futures
import os
import time

import numpy as np

from . import runtime as rt

DATA_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "finfet_fields.npz")
HELD_OUT = 64                  # the last cases of the data set are held out of training

CONFIG = "finfet_callaway.toml"
GRID = (64, 32, 32)
DT = 1.0                       # ps
FIN_BASE_NM = 50.0             # substrate height; the fin stands on it
FIN_XY_NM = (100.0, 50.0)      # fin centre in x, y
RANGES = {"height": (0.0, 100.0), "sigma": (3.0, 15.0), "power": (1e19, 1e21), "t_bg": (250.0, 400.0)}
PARAM_KEYS = ("height", "sigma", "power", "t_bg")
T_MID, T_HALF = 325.0, 75.0    # reservoir channel: (t_bg - T_MID) / T_HALF
EV_PS_PER_WATT = 6.241509074e6  # eV/ps per W
Y_UNIT_W = 1e-6                # the target is the rise per microwatt of source power
Y_SCALE = 10.0                 # K per microwatt per unit of the operator output
P_CAP_W = 5e-4                 # largest total source power a case runs at


def sample_cases(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        p = {}
        for k in PARAM_KEYS:
            lo, hi = RANGES[k]
            p[k] = float(np.exp(rng.uniform(np.log(lo), np.log(hi)))) if k == "power" else float(rng.uniform(lo, hi))
        out.append(p)
    return out


def run_power_density(p):
    """The power density a case runs at: the sampled one, or less when the
    Gaussian's total power exceeds P_CAP_W.  Linearized transport is linear in
    the source (a run at half the power has half the rise, carrier for
    carrier), so the rise per microwatt does not depend on it; far above the cap
    the linearized deviations drive occupations negative and the source
    deposit fails (as the reference's does)."""
    total = p["power"] * (2.0 * np.pi) ** 1.5 * (p["sigma"] * 1e-9) ** 3
    return p["power"] * min(1.0, P_CAP_W / total)


def case_text(p, carriers=None, seed=None, dt=DT):
    t = f"{p['t_bg']:.6f}"
    o = {"simulation.background_temperature": t, "simulation.lifetime_temperature": t,
         "simulation.initial_temperature": t, "simulation.time_step": f"{dt}",
         "boundary.boundary_values": f"[{t}, 10.0]",
         "heat_source.center": f"[{FIN_XY_NM[0]}, {FIN_XY_NM[1]}, {FIN_BASE_NM + p['height']:.6f}]",
         "heat_source.sigma": f"[{p['sigma']:.6f}, {p['sigma']:.6f}, {p['sigma']:.6f}]",
         "heat_source.power_density": f"{run_power_density(p):.6e}"}
    if carriers is not None:
        o["simulation.particle_count"] = int(carriers)
    if seed is not None:
        o["simulation.random_seed"] = int(seed)
    return rt.config_text(CONFIG, **o)


def problem_of(p, carriers=None, seed=None, dt=DT):
    return rt.Problem(case_text(p, carriers, seed, dt), input_directory=rt.CONFIGS)


class Geometry:
    """Voxel positions of the solver's cells on the operator grid (a cell is the
    voxel whose centre lies in the solid)."""

    def __init__(self, problem):
        s, m, g, i, rough, res = problem.views()
        if tuple(g.grid) != GRID:
            raise ValueError(f"the problem grid {tuple(g.grid)} is not {GRID}")
        centers, volumes = problem.cells()
        bmin, inv = np.array(g.bmin[:]), np.array(g.inv_cell[:])
        idx = np.floor((centers - bmin) * inv).astype(int)
        self.pos = (idx[:, 0] * GRID[1] + idx[:, 1]) * GRID[2] + idx[:, 2]
        if len(np.unique(self.pos)) != len(centers):
            raise ValueError("cells do not map one to one onto voxels")
        self.n_cells = len(centers)
        self.volumes = volumes
        self.mask = np.zeros(np.prod(GRID), dtype=np.float32)
        self.mask[self.pos] = 1.0

    def to_grid(self, cell_values):
        out = np.zeros(np.prod(GRID), dtype=np.float32)
        out[self.pos] = cell_values
        return out

    def to_cells(self, grid_values):
        return np.asarray(grid_values, dtype=np.float64)[self.pos]


def source_power(problem):
    """Total power of the heat source (W) and its cell rates (eV/ps)."""
    hs = problem.heat_source()
    rate = rt.arr(hs.rate, hs.n_cells)
    return float(rate.sum()) / EV_PS_PER_WATT, rate


def features(problem, p, geom):
    """Operator input [S, 3]: source density / its peak, solid mask, (t_bg - 325) / 75 in the solid."""
    _, rate = source_power(problem)
    density = rate / geom.volumes
    f = np.zeros((np.prod(GRID), 3), dtype=np.float32)
    f[geom.pos, 0] = density / density.max()
    f[geom.pos, 1] = 1.0
    f[geom.pos, 2] = (p["t_bg"] - T_MID) / T_HALF
    return f


def field_from_prediction(problem, p, geom, pred):
    """Per-cell initial temperatures from an operator output row."""
    watts, _ = source_power(problem)
    return p["t_bg"] + geom.to_cells(pred) * Y_SCALE * (watts / Y_UNIT_W)


def build_problems(cases, carriers, seed0, threads=16, dt=DT):
    """Problems of the cases, built on host threads (the mesh, its grid and the
    material tables are shared between them)."""
    first = problem_of(cases[0], carriers, seed0, dt)  # fills the caches
    with concurrent.futures.ThreadPoolExecutor(threads) as pool:
        rest = list(pool.map(lambda kp: problem_of(kp[1], carriers, seed0 + kp[0], dt), enumerate(cases[1:], 1)))
    return [first] + rest


def generate(cases, carriers, burn_in, average, sample=10, seed0=500000, dt=DT, log=print):
    """Steady cell temperatures of every case from one batch of device
    contexts: burn_in steps from t_bg, then the mean over `average` steps
    sampled every `sample`.  Returns (problems, [n_cases][n_cells] float64)."""
    t0 = time.perf_counter()
    problems = build_problems(cases, carriers, seed0, dt=dt)
    log(f"[gen] {len(cases)} problems built in {time.perf_counter() - t0:.1f} s")
    ctx = [rt.Context(p) for p in problems]
    log(f"[gen] contexts initialized at {time.perf_counter() - t0:.1f} s")
    batch = rt.Batch(ctx)
    last = time.perf_counter()
    done = 0
    while done < burn_in:  # short launches, a progress line every few seconds
        k = min(10, burn_in - done)
        batch.step(k)
        done += k
        if time.perf_counter() - last > 5.0 or done == burn_in:
            last = time.perf_counter()
            log(f"[gen] burn-in {done}/{burn_in} steps at {last - t0:.1f} s")
    acc = [np.zeros(c.n_cells) for c in ctx]
    nsamp = average // sample
    for s in range(nsamp):
        batch.step(sample)
        for k, c in enumerate(ctx):
            acc[k] += c.cells()[0]
        if time.perf_counter() - last > 5.0 or s == nsamp - 1:
            last = time.perf_counter()
            log(f"[gen] averaging {s + 1}/{nsamp} samples at {last - t0:.1f} s")
    batch.close()
    fields = np.array([a / nsamp for a in acc])
    for c in ctx:
        c.close()
    return problems, fields


def quadrant(geom):
    """Grid indices of the solid voxels of the x < nx/2, y < ny/2 quadrant and of
    their three mirror images (the fin, its substrate, the source and the
    reservoir are symmetric under x -> nx-1-x and y -> ny-1-y)."""
    X, Y, Z = GRID
    solid = geom.mask.reshape(GRID) > 0
    if not (np.array_equal(solid, solid[::-1]) and np.array_equal(solid, solid[:, ::-1])):
        raise ValueError("the cell set is not mirror-symmetric")
    x, y, z = np.nonzero(solid[: X // 2, : Y // 2])
    flat = lambda a, b: (a * Y + b) * Z + z  # noqa: E731
    return np.stack([flat(x, y), flat(X - 1 - x, y), flat(x, Y - 1 - y), flat(X - 1 - x, Y - 1 - y)])


def fold(grid_rows, quad):
    """Mirror average of full-grid rows [n, S] onto the quadrant [n, q]."""
    g = np.asarray(grid_rows, dtype=np.float64)
    return g[:, quad].mean(axis=1)


def unfold(quad_rows, quad):
    """Quadrant rows [n, q] back onto the full grid [n, S] (zero outside the solid)."""
    q = np.asarray(quad_rows)
    out = np.zeros((q.shape[0], int(np.prod(GRID))), dtype=np.float32)
    for k in range(4):
        out[:, quad[k]] = q
    return out


def pack_dataset(raw_paths, out, geometry_problem):
    """The committed data set from raw generator files (np.savez of params,
    watts, rise [n, n_cells]): each case's steady rise per microwatt of source
    power on the grid, mirror-averaged onto the quadrant (which also halves the
    Monte Carlo noise), stored as 16 bits per value with a per-case offset and
    scale."""
    geom = Geometry(geometry_problem)
    quad = quadrant(geom)
    params, g = [], []
    for path in raw_paths:
        d = np.load(path, allow_pickle=False)
        per_uw = d["rise"].astype(np.float64) / (d["watts"][:, None] / Y_UNIT_W)
        g.append(fold(np.stack([geom.to_grid(r) for r in per_uw]), quad))
        params.append(d["params"])
    g, params = np.concatenate(g), np.concatenate(params)
    lo, hi = g.min(axis=1), g.max(axis=1)
    scale = np.where(hi > lo, (hi - lo) / 65535.0, 1.0)
    q = np.clip(np.rint((g - lo[:, None]) / scale[:, None]), 0, 65535).astype(np.uint16)
    np.savez_compressed(out, params=params, g_q=q, g_lo=lo, g_scale=scale, dt=DT, config=CONFIG)


def load_dataset_cases(path):
    d = np.load(path, allow_pickle=False)
    return [dict(zip(PARAM_KEYS, map(float, row))) for row in d["params"]]


def load_dataset(path, geometry_problem):
    """(cases, rise per microwatt [n, S] on the full grid, geometry)."""
    d = np.load(path, allow_pickle=False)
    cases = [dict(zip(PARAM_KEYS, map(float, row))) for row in d["params"]]
    geom = Geometry(geometry_problem)
    g = d["g_lo"][:, None] + d["g_q"].astype(np.float64) * d["g_scale"][:, None]
    return cases, unfold(g, quadrant(geom)), geom


def training_arrays(cases, per_uw, geom, problems=None):
    """Inputs [n, S, 3], targets (rise per microwatt / Y_SCALE) [n, S] and
    weights (the solid mask) [n, S]."""
    n, S = len(cases), int(np.prod(GRID))
    x = np.zeros((n, S, 3), dtype=np.float32)
    if problems is None:
        problems = build_problems(cases, 20000, 1)
    for k, (p, pr) in enumerate(zip(cases, problems)):
        x[k] = features(pr, p, geom)
    y = (np.asarray(per_uw) / Y_SCALE).astype(np.float32)
    w = np.broadcast_to(geom.mask, (n, S)).astype(np.float32)
    return x, y, w


class WarmStart3d:
    """The FinFET operator (64 x 32 x 32, 3 channels, 4 spectral layers, width
    32, 12 modes per axis) with its training schedule."""

    def __init__(self, width=32, layers=4, modes=12, hidden=128, max_batch=4, seed=20261205):
        from . import fno
        self.op = fno.Operator3d(*GRID, in_channels=3, width=width, layers=layers, modes=modes, hidden=hidden,
                                 max_batch=max_batch, seed=seed)

    def fit(self, x, y, w, epochs=40, batch=4, lr=1e-3, stages=4, seed=1, log=None):
        """Adam on the device kernels, the rate halved at each of `stages` parts."""
        losses = []
        per = max(1, epochs // stages)
        for k in range(stages):
            t0 = time.perf_counter()
            losses.append(self.op.train(x, y, w, per, batch=batch, lr=lr * 0.5 ** k, seed=seed + k))
            if log:
                log(f"[fno] stage {k + 1}/{stages}: loss {losses[-1][-1]:.4e} ({time.perf_counter() - t0:.1f} s)")
        return np.concatenate(losses)

    def predict(self, x):
        return self.op.predict(x)

    def initial_field(self, problem, p, geom):
        """Per-cell initial temperatures: the prediction, mirror-averaged."""
        quad = quadrant(geom)
        pred = unfold(fold(self.predict(features(problem, p, geom)[None]), quad), quad)[0]
        return field_from_prediction(problem, p, geom, pred)


def relative_weights(y, mask):
    """Per-point loss weights [n, S]: the solid mask over each case's mean
    square target, so that every case counts by its relative error."""
    ms = (y.astype(np.float64) ** 2 * mask).sum(axis=1) / mask.sum(axis=1)
    return (mask / np.maximum(ms, 1e-30)[:, None]).astype(np.float32)


def train(epochs=48, lr=1e-3, stages=4, path=DATA_PATH, log=None):
    """The operator trained on the data set but its last HELD_OUT cases.
    Returns (model, info) with the held-out relative RMS error of the rise
    over the solid cells, the training time and the loss history."""
    cases_all = load_dataset_cases(path)
    geometry_problem = problem_of(cases_all[0], 20000, 1)
    cases, per_uw, geom = load_dataset(path, geometry_problem)
    x, y, mask = training_arrays(cases, per_uw, geom)
    w = relative_weights(y, mask)
    n = len(cases) - HELD_OUT
    model = WarmStart3d()
    t0 = time.perf_counter()
    losses = model.fit(x[:n], y[:n], w[:n], epochs=epochs, lr=lr, stages=stages, log=log)
    seconds = time.perf_counter() - t0
    quad = quadrant(geom)
    pred = np.concatenate([unfold(fold(model.predict(x[k:k + 4]), quad), quad) for k in range(n, len(cases), 4)])
    err = np.sqrt(((pred - y[n:]) ** 2 * mask[n:]).sum(axis=1) / (y[n:] ** 2 * mask[n:]).sum(axis=1))
    info = {"train_cases": n, "held_out_cases": len(cases) - n, "epochs": int(len(losses)),
            "train_seconds": seconds, "final_loss": float(losses[-1]),
            "held_out_relative_rms": float(err.mean()), "held_out_relative_rms_max": float(err.max())}
    if log:
        log(f"[fno] trained on {n} cases in {seconds:.1f} s; held-out relative RMS {err.mean():.4f} "
            f"(max {err.max():.4f})")
    return model, info, (x, y, w)


def gemm_roofline(model, x, y, w, reps=10, peak_tflops=2500.0):
    """Roofline object of the operator's GEMMs: forward + backward passes at
    the training batch, algorithmic flops (2 m n k per product) over the
    HIP-event durations of the GEMM launches."""
    op = model.op
    b = op.shape.max_batch
    op.set_timing(True)
    op.gemm_stats(reset=True)
    for _ in range(reps):
        op.loss_grad(x[:b], y[:b], w[:b])
    flops, launches, ms = op.gemm_stats(reset=True)
    op.set_timing(False)
    tflops = flops / (ms / 1e3) / 1e12
    return {"bound": "mfma", "achieved": tflops, "peak": peak_tflops, "unit": "TFLOP/s", "frac": tflops / peak_tflops,
            "traffic": None, "kernel": "k_gemm", "flops_per_pass": flops / reps, "launches_per_pass": launches / reps,
            "gemm_ms_per_pass": ms / reps, "avg_launch_ms": ms / max(1, launches)}


def run_windows(problem, steps, window, sample, initial_field=None, seed=None):
    """Window averages [steps // window, n_cells] of a run and the wall seconds at
    the end of each window (from context creation on)."""
    t0 = time.perf_counter()
    ctx = rt.Context(problem, seed=seed, initial_field=initial_field)
    try:
        nwin = steps // window
        avg = np.zeros((nwin, ctx.n_cells))
        wall = np.zeros(nwin)
        per = window // sample
        for k in range(nwin):
            acc = np.zeros(ctx.n_cells)
            for _ in range(per):
                ctx.step(sample)
                acc += ctx.cells()[0]
            avg[k] = acc / per
            wall[k] = time.perf_counter() - t0
        return avg, wall
    finally:
        ctx.close()


def first_below(err, tol):
    hit = np.nonzero(np.asarray(err) <= tol)[0]
    return int(hit[0]) if len(hit) else None


def reference_fields(cases, carriers, burn_in, average, sample=10, seed0=900000, log=None):
    """T_ref of each case: the mean of a plain run (its own seed) over `average`
    steps after `burn_in`, the cases stepped together as one batch."""
    problems = build_problems(cases, carriers, seed0)
    ctx = [rt.Context(pr) for pr in problems]
    batch = rt.Batch(ctx)
    try:
        done = 0
        while done < burn_in:
            k = min(500, burn_in - done)
            batch.step(k)
            done += k
            if log:
                log(f"[ref] burn-in {done}/{burn_in}")
        acc = [np.zeros(c.n_cells) for c in ctx]
        for _ in range(average // sample):
            batch.step(sample)
            for a, c in zip(acc, ctx):
                a += c.cells()[0]
        return [a / (average // sample) for a in acc]
    finally:
        batch.close()
        for c in ctx:
            c.close()


def evaluate(model, p, carriers, t_ref, steps, window=20, sample=5, tol=0.5, seed=4242, log=None):
    """Plain against warm-started Monte Carlo on one case, with common random
    numbers.  Three contexts share the seed, and therefore every carrier
    position, mode and scattering draw: the transport is linear in the
    occupations, so runs that start from different fields differ only by the
    decaying image of their initial difference (the stationary noise is common
    to all three).  The yardstick starts at T_ref (a long plain run's mean),
    the plain run at t_bg, the warm run at the operator's field.  A run is
    steady at the first window whose average is within `tol` K RMS of the
    yardstick's same window; the plain and warm runs are timed by the wall
    clock, each from context creation, the operator's prediction included."""
    problem = problem_of(p, carriers, seed)
    geom = Geometry(problem)
    yard, _ = run_windows(problem, steps, window, sample, initial_field=t_ref, seed=seed)
    plain, plain_wall = run_windows(problem, steps, window, sample, seed=seed)
    t0 = time.perf_counter()
    field = model.initial_field(problem, p, geom)
    t_pred = time.perf_counter() - t0
    warm, warm_wall = run_windows(problem, steps, window, sample, initial_field=field, seed=seed)
    err_plain = np.sqrt(np.mean((plain - yard) ** 2, axis=1))
    err_warm = np.sqrt(np.mean((warm - yard) ** 2, axis=1))
    k_plain, k_warm = first_below(err_plain, tol), first_below(err_warm, tol)
    rise = t_ref - p["t_bg"]
    out = {
        "case": p, "carriers": carriers, "window_steps": window, "tol_k": tol,
        "steady_rms_rise_k": float(np.sqrt(np.mean(rise ** 2))), "peak_rise_k": float(rise.max()),
        "operator_rms_k": float(np.sqrt(np.mean((field - t_ref) ** 2))),
        "operator_rel_rms": float(np.sqrt(np.mean((field - t_ref) ** 2) / np.mean(rise ** 2))),
        "plain_steps_to_steady": None if k_plain is None else (k_plain + 1) * window,
        "warm_steps_to_steady": None if k_warm is None else (k_warm + 1) * window,
        "plain_seconds": None if k_plain is None else float(plain_wall[k_plain]),
        "warm_seconds": None if k_warm is None else float(warm_wall[k_warm] + t_pred),
        "predict_seconds": t_pred,
        "err_plain_k": err_plain.tolist(), "err_warm_k": err_warm.tolist(),
    }
    if k_plain is not None and k_warm is not None:
        out["speedup_steps"] = (k_plain + 1) / (k_warm + 1)
        out["speedup_wall"] = out["plain_seconds"] / out["warm_seconds"]
    if log:
        log(f"[eval] RMS rise {out['steady_rms_rise_k']:.1f} K (peak {out['peak_rise_k']:.0f} K); operator "
            f"{out['operator_rms_k']:.2f} K RMS ({100 * out['operator_rel_rms']:.2f} %); {tol} K RMS after "
            f"{out['plain_steps_to_steady']} plain vs {out['warm_steps_to_steady']} warm steps; "
            f"{out['plain_seconds']} vs {out['warm_seconds']} s")
    return out


def evaluate_batch(model, cases, carriers, t_refs, steps, window=20, sample=5, tol=0.5, seed0=700000, log=None):
    """evaluate() over many cases as one batch (yardstick, plain and warm
    contexts of every case stepped together): steps to steady per case."""
    problems = build_problems(cases, carriers, seed0)
    geom = Geometry(problems[0])
    fields = [model.initial_field(pr, p, geom) for pr, p in zip(problems, cases)]
    ctx = []
    for k, pr in enumerate(problems):
        ctx += [rt.Context(pr, initial_field=t_refs[k]), rt.Context(pr), rt.Context(pr, initial_field=fields[k])]
    batch = rt.Batch(ctx)
    nwin = steps // window
    err = np.zeros((len(cases), 2, nwin))
    try:
        for w in range(nwin):
            acc = [np.zeros(c.n_cells) for c in ctx]
            for _ in range(window // sample):
                batch.step(sample)
                for a, c in zip(acc, ctx):
                    a += c.cells()[0]
            for k in range(len(cases)):
                y = acc[3 * k]
                err[k, 0, w] = np.sqrt(np.mean((acc[3 * k + 1] - y) ** 2)) / (window // sample)
                err[k, 1, w] = np.sqrt(np.mean((acc[3 * k + 2] - y) ** 2)) / (window // sample)
            if log and (w + 1) % 25 == 0:
                log(f"[eval] {(w + 1) * window}/{steps} steps")
    finally:
        batch.close()
        for c in ctx:
            c.close()
    rows = []
    for k, p in enumerate(cases):
        rise = t_refs[k] - p["t_bg"]
        kp, kw = first_below(err[k, 0], tol), first_below(err[k, 1], tol)
        rows.append({"case": p, "steady_rms_rise_k": float(np.sqrt(np.mean(rise ** 2))),
                     "operator_rms_k": float(np.sqrt(np.mean((fields[k] - t_refs[k]) ** 2))),
                     "operator_rel_rms": float(np.sqrt(np.mean((fields[k] - t_refs[k]) ** 2) / np.mean(rise ** 2))),
                     "plain_steps_to_steady": None if kp is None else (kp + 1) * window,
                     "warm_steps_to_steady": None if kw is None else (kw + 1) * window})
    return rows, err


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="Steady FinFET fields from the batched device step.")
    ap.add_argument("--cases", type=int, default=768)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--carriers", type=int, default=250000)
    ap.add_argument("--burn-in", type=int, default=3000)
    ap.add_argument("--average", type=int, default=1000)
    ap.add_argument("--first", type=int, default=0, help="index of the first case of the sample to run")
    ap.add_argument("--evaluate", type=int, default=0, metavar="N",
                    help="instead: train on the committed data set and evaluate N fresh held-out cases")
    ap.add_argument("--epochs", type=int, default=48)
    ap.add_argument("--steps", type=int, default=3000, help="evaluation: steps of the plain and warm runs")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)  # noqa: E731
    if a.evaluate:
        import json
        model, info, _ = train(epochs=a.epochs, log=log)
        cases = sample_cases(a.evaluate, a.seed + 1000)  # parameter sets outside the data set
        t0 = time.perf_counter()
        t_refs = reference_fields(cases, a.carriers, a.burn_in, a.average, log=log)
        log(f"[eval] reference fields in {time.perf_counter() - t0:.1f} s")
        rows, err = evaluate_batch(model, cases, a.carriers, t_refs, a.steps, log=log)
        sp = [r["plain_steps_to_steady"] / r["warm_steps_to_steady"] for r in rows
              if r["plain_steps_to_steady"] and r["warm_steps_to_steady"]]
        summary = {"operator": info, "carriers": a.carriers, "steps": a.steps, "reference_burn_in": a.burn_in,
                   "reference_average": a.average, "cases": rows,
                   "speedup_steps_median": float(np.median(sp)) if sp else None,
                   "speedup_steps_geomean": float(np.exp(np.mean(np.log(sp)))) if sp else None,
                   "reached": len(sp), "evaluated": len(rows)}
        with open(a.out, "w") as f:
            json.dump(summary, f)
        np.save(a.out.replace(".json", "_err.npy"), err.astype(np.float32))
        log(f"[eval] {len(sp)}/{len(rows)} cases reached 0.5 K RMS in both runs; step speed-up median "
            f"{summary['speedup_steps_median']}, geometric mean {summary['speedup_steps_geomean']}")
        return
    cases = sample_cases(a.first + a.cases, a.seed)[a.first:]
    problems, fields = generate(cases, a.carriers, a.burn_in, a.average, seed0=500000 + a.first, log=log)
    rises = fields - np.array([c["t_bg"] for c in cases])[:, None]
    watts = np.array([source_power(p)[0] for p in problems])
    np.savez(a.out, params=np.array([[c[k] for k in PARAM_KEYS] for c in cases]), watts=watts,
             rise=rises.astype(np.float16), seed=a.seed, first=a.first, carriers=a.carriers, burn_in=a.burn_in,
             average=a.average, dt=DT)
    log(f"[gen] wrote {a.out}: {len(cases)} cases, RMS rise per case {np.sqrt((rises ** 2).mean(axis=1)).min():.3g}"
        f" .. {np.sqrt((rises ** 2).mean(axis=1)).max():.3g} K")


if __name__ == "__main__":
    main()
