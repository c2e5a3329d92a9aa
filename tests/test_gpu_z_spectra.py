"""Weighted and Gaussian heat-source spectra on the device.

- Carrier by carrier against the oracle over 3 steps (c1sw: branch weights;
  c1sg: 5 THz Gaussian of 1.5 THz; and a 0.05 THz Gaussian at 40 THz, above every
  mode, whose weights survive only by the per-cell peak normalization):
  identical carrier sets, modes and cells, occupations to 1e-12, deposited
  energy equal to the budget.
- Whole runs against 16 runs of the reference (tests/golden/run_c1sw.npz,
  run_c1sg.npz) with the z-score rules of test_gpu_statistics.
"""
import os

import numpy as np
import pytest

from oracle import oracle_ctypes as O
from phonomc_amd import runtime as rt

from test_gpu_statistics import WORKLOADS, gpu_runs, z_score

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
MASK = (1 << 28) - 1
WORKLOADS.update({"c1sw": "c1sw_weighted.toml", "c1sg": "c1sg_gaussian.toml"})
NARROW = {"heat_source.frequency_center": "40.0", "heat_source.frequency_sigma": "0.05"}


def by_uid(c):
    live = c["mode"] >= 0
    order = np.argsort(c["uid"][live])
    return {k: v[live][order] for k, v in c.items()}


@pytest.mark.parametrize("cfg,over", [("c1sw_weighted.toml", {}), ("c1sg_gaussian.toml", {}),
                                      ("c1sg_gaussian.toml", NARROW)])
def test_source_spectra_step_like_the_oracle(cfg, over):
    print(f"case {cfg} {over} n=20000", flush=True)
    p = rt.Problem(rt.config_text(cfg, **{"simulation.particle_count": 20000, **over}), input_directory=rt.CONFIGS,
                   mode_sampling=1)
    g = rt.Context(p)
    o = O.Sim(p)
    o.initialize()
    for step in range(3):
        g.step(1)
        o.step(1)
        a, b = by_uid(g.carriers()), o.carriers()
        assert np.array_equal(a["uid"], b["uid"]), f"step {step}: live carrier sets differ"
        assert np.array_equal(a["mode"] & MASK, b["mode"]), f"step {step}: modes differ"
        assert np.array_equal(a["cell"], b["cell"]), f"step {step}: cells differ"
        np.testing.assert_allclose(a["occupation"], b["occupation"], rtol=1e-12, atol=1e-300)
        lg, lo = g.ledger(), o.ledger()
        for k in ["source_updates", "step_source_updates", "step_absorbed", "step_injected"]:
            assert lg[k] == lo[k], k
        for k in ["step_source_energy_ev", "thermal_energy_ev", "source_pending_ev"]:
            assert lg[k] == pytest.approx(lo[k], rel=1e-11, abs=1e-9), k
        assert abs(lg["step_balance_residual_ev"]) <= 1e-12 * lg["thermal_energy_ev"]
    assert g.ledger()["source_updates"] > 0


@pytest.mark.parametrize("key", ["c1sw", "c1sg"])
def test_whole_run_statistics_match_reference(key):
    golden = np.load(os.path.join(HERE, "golden", f"run_{key}.npz"))
    print(f"case run_{key} n={int(golden['n'])} steps={int(golden['steps'])}", flush=True)
    temps, keff, kint, _ = gpu_runs(key, golden)
    zt = z_score(temps, golden["temperature"])
    spread = (temps.std(axis=0) > 0) | (golden["temperature"].std(axis=0) > 0)
    print(f"{key}: cell T max|z| {np.abs(zt).max():.2f}, mean z^2 {np.mean(zt[spread] ** 2):.2f}")
    assert np.abs(zt).max() <= 4.5
    assert np.mean(zt[spread] ** 2) <= 2.5
