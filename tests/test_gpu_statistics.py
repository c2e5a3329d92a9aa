"""Whole-run statistics of the device step against the reference solver.

tests/golden/run_c1.npz, run_c2.npz, run_c1s.npz, run_c1c.npz, run_c2c.npz, run_c1cs.npz, run_finfet.npz,
run_c1l.npz and run_finfet_cw.npz hold 16 runs of the
reference PhonoMC build per workload (seeds 1000..1015, written by tests/golden/make_golden.py):
second-half averages of every cell temperature and of kappa_int / kappa_eff as
the reference writes them to convergence.txt.  The same TOML inputs and seeds
run here through the C++ facade on the GPU.  The random streams differ (Philox
here, mt19937_64 in the reference), so the comparison is distributional: for
each quantity z = (mean_gpu - mean_ref) / sqrt(se_gpu^2 + se_ref^2) over the 16
seeds, with |z| <= 3 for kappa and, for the cell temperatures, max |z| <= 4.5
(a 3-sigma bound corrected for up to 500 cells) and mean z^2 <= 2.5.  For the
layered film the interface event and transmission totals of each run (the
reference's MonteCarloSolver::interface_event_count and
interface_transmitted_count) are compared the same way, |z| <= 3.

The second test states the error of the FP32 flight against the all-FP64
oracle: both start from the same population and run the same Philox streams.
"""
import os
import tempfile

import numpy as np
import pytest

from phonomc_amd import runtime as rt

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
WORKLOADS = {"c1": "c1_cross_plane.toml", "c2": "c2_inplane_film.toml", "c1s": "c1s_source.toml",
             "c1c": "c1c_callaway.toml", "c2c": "c2c_callaway.toml", "finfet": "finfet.toml",
             "c1cs": "c1cs_callaway_source.toml", "c1l": "c1l_layered.toml",
             "finfet_cw": "finfet_callaway.toml"}
# Inputs the reference runs were made with beyond the bundled file (tests/golden/make_golden.py).
EXTRA = {"finfet_cw": {"simulation.grid_xyz": "[20, 10, 15]"}}


def second_half(folder, ncell, steps):
    """Averages of the convergence rows of the second half of a run (the rule of make_golden.py)."""
    rows = [line.split() for line in open(os.path.join(folder, "convergence.txt")) if not line.startswith("#")]
    half = [r for r in rows if int(r[0]) > steps // 2]
    t = np.array([[float(x) for x in r[2:2 + ncell]] for r in half]).mean(axis=0)
    kint = np.mean([float(r[2 + ncell + 1]) for r in half])
    keff = np.mean([float(r[2 + ncell + 2]) for r in half])
    return t, keff, kint


def summary_counts(path):
    """events_total / transmitted_total of the [material_interface_diagnostics] section."""
    text = open(path).read()
    section = text[text.index("[material_interface_diagnostics]"):]
    get = lambda key: int(section.split(f"\n{key} = ")[1].split()[0])
    return get("events_total"), get("transmitted_total")


def gpu_runs(key, golden):
    n, steps, interval = int(golden["n"]), int(golden["steps"]), int(golden["interval"])
    temps, keff, kint, iface = [], [], [], []
    for seed in golden["seeds"]:
        tmp = tempfile.mkdtemp(prefix=f"pmc_{key}_")
        text = rt.config_text(WORKLOADS[key], **{"simulation.particle_count": n, "simulation.iterations": steps,
                                                 "simulation.random_seed": int(seed),
                                                 "simulation.convergence_write_interval": interval,
                                                 "io.output_folder": f'"{tmp}/out"', **EXTRA.get(key, {})})
        solver = rt.Solver(rt.Problem(text, input_directory=rt.CONFIGS), os.path.join(tmp, "out"))
        solver.run(steps)
        ncell = solver.n_cells
        if "interface" in golden:
            solver.append_summary(os.path.join(tmp, "summary.txt"))
            iface.append(summary_counts(os.path.join(tmp, "summary.txt")))
        solver.close()
        t, ke, ki = second_half(os.path.join(tmp, "out"), ncell, steps)
        temps.append(t)
        keff.append(ke)
        kint.append(ki)
    return np.array(temps), np.array(keff), np.array(kint), np.array(iface)


def z_score(a, b):
    """Per-quantity z over the seeds.  A quantity with no spread in either set
    (a cell the heat has not reached keeps T_bg exactly) scores 0 when the two
    means agree and inf when they do not."""
    se = np.sqrt(a.var(axis=0, ddof=1) / len(a) + b.var(axis=0, ddof=1) / len(b))
    diff = a.mean(axis=0) - b.mean(axis=0)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(se > 0, diff / np.where(se > 0, se, 1.0), np.where(diff == 0, 0.0, np.inf))


@pytest.mark.parametrize("key", ["c1", "c2", "c1s", "c1c", "c2c", "c1cs", "finfet", "c1l", "finfet_cw"])
def test_whole_run_statistics_match_reference(key):
    """c1s adds the Gaussian heat source (kappa is not estimated with a source);
    c1c and c2c run linearized Callaway scattering (c2c with strict-elastic rough walls);
    c1cs adds the heat source to c1c (deposits by mode heat capacity); finfet runs the
    4080-triangle FinFET mesh with a heat source in the fin; c1l is the Si/Ge
    layered film with a diffuse-mismatch interface; finfet_cw is config 3's
    physics (Callaway N/U + isotope on the FinFET, strict-elastic diffuse walls,
    1e20 W/m^3 source, fixed-flux reservoir) on a 20 x 10 x 15 grid."""
    golden = np.load(os.path.join(HERE, "golden", f"run_{key}.npz"))
    temps, keff, kint, iface = gpu_runs(key, golden)
    zt = z_score(temps, golden["temperature"])
    spread = (temps.std(axis=0) > 0) | (golden["temperature"].std(axis=0) > 0)
    msg = (f"{key}: cell T max|z| {np.abs(zt).max():.2f}, mean z^2 {np.mean(zt[spread] ** 2):.2f} over "
           f"{spread.sum()} cells with spread ({(~spread).sum()} cells at T_bg in every run)")
    if np.any(golden["kappa_eff"] != 0.0):
        zke = z_score(keff, golden["kappa_eff"])
        zki = z_score(kint, golden["kappa_int"])
        msg += (f"; kappa_eff {keff.mean():.3f} vs {golden['kappa_eff'].mean():.3f} (z {zke:+.2f}); "
                f"kappa_int {kint.mean():.3f} vs {golden['kappa_int'].mean():.3f} (z {zki:+.2f})")
        assert abs(zke) <= 3.0 and abs(zki) <= 3.0
    if "interface" in golden:
        zi = z_score(iface.astype(float), golden["interface"].astype(float))
        msg += (f"; interface events {iface[:, 0].mean():.1f} vs {golden['interface'][:, 0].mean():.1f} (z {zi[0]:+.2f}),"
                f" transmitted {iface[:, 1].mean():.1f} vs {golden['interface'][:, 1].mean():.1f} (z {zi[1]:+.2f})")
        assert np.all(np.abs(zi) <= 3.0)
    print(msg)
    assert np.abs(zt).max() <= 4.5
    assert np.mean(zt[spread] ** 2) <= 2.5


def test_fp32_flight_error_against_fp64_oracle():
    """Cross-plane film, 20 000 carriers, 100 steps from one population: cell
    temperatures and the per-cell heat flux of the FP32-flight device path
    against the all-FP64 oracle.  The difference is stated and bounded well
    below the statistical error of such a run (about 2e-3 K per cell)."""
    from oracle import oracle_ctypes as O

    p = rt.Problem(rt.config_text("c1_cross_plane.toml", **{"simulation.particle_count": 20000}), mode_sampling=1)
    g = rt.Context(p)
    o = O.Sim(p)
    o.initialize()
    g.step(100)
    o.step(100)
    tg, _ = g.cells()
    to, _ = o.cells()
    qg, qo = g.heat_flux()[:, 2], o.heat_flux()[:, 2]
    dt = np.abs(tg - to).max()
    dq = np.abs(qg.mean() - qo.mean()) / np.abs(qo.mean())
    print(f"FP32 flight vs FP64 oracle after 100 steps: max |dT| = {dt:.3e} K, mean flux rel. diff = {dq:.3e}")
    assert dt < 2e-4
    assert dq < 1e-2
