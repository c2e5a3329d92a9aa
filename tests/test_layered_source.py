"""Heat sources in layered materials (c1ls_source.toml: the Si/Ge film of c1l with a
Gaussian source on the interface).  PhononSourceOperator configures every
material (src/solver/PhononSourceOperator.cpp:20-103) and deposits into each
cell through the cell's material (:121-230).

- The host's per-mode source arrays of a layered problem are the single-material
  arrays of each material, concatenated in the order of the global active bank,
  for the thermal, weighted and Gaussian spectra (the single-material arrays are
  pinned to the reference by test_oracle_golden's source fixtures).
- The oracle's energy balance closes every step with the source running in both
  materials, and the deposited energy equals the prescribed budget of the cells
  that hold eligible carriers.
- The oracle's whole runs agree with 16 reference runs (tests/golden/run_c1ls.npz,
  written by make_golden.py --layered-source): cell temperatures and interface
  totals by the z-score rules of test_oracle_statistics.
"""
import concurrent.futures
import os

import numpy as np
import pytest

from oracle import oracle_ctypes as O
from phonomc_amd import runtime as rt

from test_oracle_statistics import oracle_run, z_score

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTED = {"heat_source.spectrum": '"weighted"', "heat_source.branch_weights": "[1.0, 0.5, 2.0, 0.0, 0.3, 0.1]"}
GAUSS = {"heat_source.spectrum": '"gaussian"', "heat_source.frequency_center": "5.0",
         "heat_source.frequency_sigma": "1.5"}


def source_arrays(p):
    h = p.heat_source()
    s, m, *_ = p.views()
    elig = rt.arr(h.mode_eligible, m.n_active, np.int32)
    weight = rt.arr(h.mode_weight, m.n_active) if h.spectrum else np.zeros(m.n_active)
    return elig, weight


@pytest.mark.parametrize("over", [{}, WEIGHTED, GAUSS])
def test_source_arrays_concatenate_the_materials(over):
    layered = rt.Problem(rt.config_text("c1ls_source.toml", **over), input_directory=rt.CONFIGS)
    elig, weight = source_arrays(layered)
    parts_e, parts_w = [], []
    for folder in ("synthetic:si16", "synthetic:ge16"):
        single = rt.Problem(rt.config_text("c1s_source.toml", **{"io.material_folder": f'"{folder}"', **over}),
                            input_directory=rt.CONFIGS)
        e, w = source_arrays(single)
        parts_e.append(e)
        parts_w.append(w)
    assert np.array_equal(elig, np.concatenate(parts_e))
    assert np.array_equal(weight, np.concatenate(parts_w))
    assert elig.sum() > 0


def test_energy_balance_and_budget():
    p = rt.Problem(rt.config_text("c1ls_source.toml", **{"simulation.particle_count": 20000}),
                   input_directory=rt.CONFIGS)
    sim = O.Sim(p, threads=1)
    sim.initialize()
    deposited = 0.0
    for _ in range(20):
        sim.step(1)
        led = sim.ledger()
        assert abs(led["step_balance_residual_ev"]) <= 1e-12 * led["thermal_energy_ev"]
        deposited += led["step_source_energy_ev"]
    led = sim.ledger()
    # every heated cell holds eligible carriers: nothing stays pending, and the
    # deposit is the prescribed power times the elapsed time
    prescribed = sum(rt.arr(p.heat_source().rate, p.heat_source().n_cells)) * 20 * 1.0
    assert led["source_pending_ev"] == pytest.approx(0.0, abs=1e-9 * prescribed)
    assert deposited == pytest.approx(prescribed, rel=1e-9)
    # both materials' cells warm
    t = sim.cells()[0]
    assert t[9] > 301.0 and t[10] > 300.5
    sim.close()


def test_oracle_runs_match_reference():
    golden = np.load(os.path.join(HERE, "golden", "run_c1ls.npz"))
    n, steps, interval = int(golden["n"]), int(golden["steps"]), int(golden["interval"])
    seeds = [8100 + k for k in range(8)]
    with concurrent.futures.ThreadPoolExecutor(8) as pool:
        runs = list(pool.map(lambda s: oracle_run("c1ls_source.toml", n, steps, interval, s, counters=True), seeds))
    temps = np.array([r[0] for r in runs])
    iface = np.array([r[1] for r in runs], dtype=float)
    z = z_score(temps, golden["temperature"])
    zi = z_score(iface, golden["interface"].astype(float))
    print(f"c1ls oracle: cell T max|z| {np.abs(z).max():.2f}, mean z^2 {np.mean(z ** 2):.2f}; interface events "
          f"z {zi[0]:+.2f}, transmitted z {zi[1]:+.2f}")
    assert np.abs(z).max() <= 4.5
    assert np.mean(z ** 2) <= 2.5
    assert np.all(np.abs(zi) <= 3.0)
