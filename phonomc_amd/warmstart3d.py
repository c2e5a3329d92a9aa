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
import time

import numpy as np

from . import runtime as rt

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


def case_text(p, carriers=None, seed=None, dt=DT):
    t = f"{p['t_bg']:.6f}"
    o = {"simulation.background_temperature": t, "simulation.lifetime_temperature": t,
         "simulation.initial_temperature": t, "simulation.time_step": f"{dt}",
         "boundary.boundary_values": f"[{t}, 10.0]",
         "heat_source.center": f"[{FIN_XY_NM[0]}, {FIN_XY_NM[1]}, {FIN_BASE_NM + p['height']:.6f}]",
         "heat_source.sigma": f"[{p['sigma']:.6f}, {p['sigma']:.6f}, {p['sigma']:.6f}]",
         "heat_source.power_density": f"{p['power']:.6e}"}
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


def target(problem, p, geom, cell_t):
    """Operator target on the grid: (T - t_bg) per microwatt, in units of Y_SCALE."""
    watts, _ = source_power(problem)
    return geom.to_grid((cell_t - p["t_bg"]) / (watts / Y_UNIT_W) / Y_SCALE)


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
    done = 0
    while done < burn_in:
        k = min(500, burn_in - done)
        batch.step(k)
        done += k
        log(f"[gen] burn-in {done}/{burn_in} steps at {time.perf_counter() - t0:.1f} s")
    acc = [np.zeros(c.n_cells) for c in ctx]
    nsamp = average // sample
    for s in range(nsamp):
        batch.step(sample)
        for k, c in enumerate(ctx):
            acc[k] += c.cells()[0]
        if s % 50 == 49:
            log(f"[gen] averaging {s + 1}/{nsamp} samples at {time.perf_counter() - t0:.1f} s")
    batch.close()
    fields = np.array([a / nsamp for a in acc])
    for c in ctx:
        c.close()
    return problems, fields


def quantize(rows):
    """Rows of floats as uint8 with a per-row offset and scale (8 bits of each row's range)."""
    lo = rows.min(axis=1)
    hi = rows.max(axis=1)
    scale = np.where(hi > lo, (hi - lo) / 255.0, 1.0)
    q = np.clip(np.rint((rows - lo[:, None]) / scale[:, None]), 0, 255).astype(np.uint8)
    return q, lo, scale


def dequantize(q, lo, scale):
    return lo[:, None] + q.astype(np.float64) * scale[:, None]


def save_dataset(path, cases, rises, meta):
    """rises: [n][n_cells] steady T - t_bg (K) in the solver's cell order."""
    q, lo, scale = quantize(np.asarray(rises))
    params = np.array([[c[k] for k in PARAM_KEYS] for c in cases])
    np.savez_compressed(path, params=params, rise_q=q, rise_lo=lo, rise_scale=scale,
                        **{k: np.asarray(v) for k, v in meta.items()})


def load_dataset(path):
    d = np.load(path, allow_pickle=False)
    cases = [dict(zip(PARAM_KEYS, map(float, row))) for row in d["params"]]
    return cases, dequantize(d["rise_q"], d["rise_lo"], d["rise_scale"]), d


def training_arrays(cases, rises, geometry_problem):
    """Inputs [n, S, 3], targets [n, S] and weights (the solid mask) [n, S] of a
    data set; features and power scales from each case's own problem (cheap:
    the geometry caches are warm)."""
    geom = Geometry(geometry_problem)
    n, S = len(cases), int(np.prod(GRID))
    x = np.zeros((n, S, 3), dtype=np.float32)
    y = np.zeros((n, S), dtype=np.float32)
    problems = build_problems(cases, 20000, 1)
    for k, (p, pr) in enumerate(zip(cases, problems)):
        x[k] = features(pr, p, geom)
        y[k] = target(pr, p, geom, p["t_bg"] + rises[k])
    w = np.broadcast_to(geom.mask, (n, S)).astype(np.float32)
    return x, y, w, geom


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
            losses.append(self.op.train(x, y, w, per, batch=batch, lr=lr * 0.5 ** k, seed=seed + k))
            if log:
                log(f"[fno] stage {k + 1}/{stages}: loss {losses[-1][-1]:.4e}")
        return np.concatenate(losses)

    def predict(self, x):
        return self.op.predict(x)

    def initial_field(self, problem, p, geom):
        pred = self.predict(features(problem, p, geom)[None])[0]
        return field_from_prediction(problem, p, geom, pred)


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


def time_to_steady(avg, t_ref, tol=0.5):
    rms = np.sqrt(np.mean((avg - t_ref[None, :]) ** 2, axis=1))
    hit = np.nonzero(rms <= tol)[0]
    return (int(hit[0]) if len(hit) else None), rms


def evaluate(model, p, carriers, ref_burn_in, ref_average, run_steps, window, sample=10, tol=0.5, seed=4242,
             log=None):
    """Plain against warm-started Monte Carlo on one case at matched statistical
    error: the same carriers, windows and sampling.  T_ref is the mean of a long
    plain run after `ref_burn_in` steps; the plain run starts from t_bg, the
    warm run from the operator's field (prediction time counted)."""
    problem = problem_of(p, carriers, seed)
    geom = Geometry(problem)
    ref, _ = run_windows(problem, ref_burn_in + ref_average, window, sample, seed=seed)
    t_ref = ref[ref_burn_in // window:].mean(axis=0)
    plain, plain_wall = run_windows(problem, run_steps, window, sample, seed=seed + 1)
    t0 = time.perf_counter()
    field = model.initial_field(problem, p, geom)
    t_pred = time.perf_counter() - t0
    warm, warm_wall = run_windows(problem, run_steps, window, sample, initial_field=field, seed=seed + 2)
    k_plain, rms_plain = time_to_steady(plain, t_ref, tol)
    k_warm, rms_warm = time_to_steady(warm, t_ref, tol)
    half = len(plain) // 2

    def rms(a):
        return float(np.sqrt(np.mean((a - t_ref) ** 2)))

    out = {
        "case": p, "carriers": carriers, "window_steps": window, "tol_k": tol,
        "steady_rms_rise_k": float(np.sqrt(np.mean((t_ref - p["t_bg"]) ** 2))),
        "peak_rise_k": float((t_ref - p["t_bg"]).max()),
        "plain_windows": None if k_plain is None else k_plain + 1,
        "warm_windows": None if k_warm is None else k_warm + 1,
        "plain_seconds": None if k_plain is None else float(plain_wall[k_plain]),
        "warm_seconds": None if k_warm is None else float(warm_wall[k_warm] + t_pred),
        "predict_seconds": t_pred,
        "operator_rms_k": rms(field),
        "uniform_rms_k": rms(np.full_like(t_ref, p["t_bg"])),
        "window_noise_rms_k": float(np.sqrt(np.mean((ref[ref_burn_in // window:] - t_ref) ** 2))),
        "plain_late_rms_k": rms(plain[half:].mean(axis=0)),
        "warm_late_rms_k": rms(warm[half:].mean(axis=0)),
        "rms_plain": rms_plain.tolist(), "rms_warm": rms_warm.tolist(),
    }
    if out["plain_windows"] and out["warm_windows"]:
        out["speedup_steps"] = out["plain_windows"] / out["warm_windows"]
        out["speedup_wall"] = out["plain_seconds"] / out["warm_seconds"]
    if log:
        log(f"[eval] steady RMS rise {out['steady_rms_rise_k']:.2f} K (peak {out['peak_rise_k']:.1f} K): operator "
            f"{out['operator_rms_k']:.3f} K RMS from T_ref (t_bg start {out['uniform_rms_k']:.2f} K); 0.5 K RMS "
            f"after {out['plain_windows']} plain vs {out['warm_windows']} warm windows of {window} steps; "
            f"second-half RMS {out['plain_late_rms_k']:.3f} K plain, {out['warm_late_rms_k']:.3f} K warm")
    return out


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="Steady FinFET fields from the batched device step.")
    ap.add_argument("--cases", type=int, default=768)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--carriers", type=int, default=250000)
    ap.add_argument("--burn-in", type=int, default=4000)
    ap.add_argument("--average", type=int, default=1000)
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    cases = sample_cases(a.cases, a.seed)
    log = lambda m: print(m, flush=True)  # noqa: E731
    problems, fields = generate(cases, a.carriers, a.burn_in, a.average, log=log)
    rises = fields - np.array([c["t_bg"] for c in cases])[:, None]
    save_dataset(a.out, cases, rises, dict(seed=a.seed, carriers=a.carriers, burn_in=a.burn_in,
                                           average=a.average, dt=DT, config=CONFIG))
    log(f"[gen] wrote {a.out}: {len(cases)} cases, RMS rise per case {np.sqrt((rises ** 2).mean(axis=1)).min():.3g}"
        f" .. {np.sqrt((rises ** 2).mean(axis=1)).max():.3g} K")


if __name__ == "__main__":
    main()
