"""Heat sources in layered materials on the device (c1ls_source.toml).

- Carrier by carrier against the oracle over 3 steps, thermal, weighted and
  Gaussian spectra: identical carrier sets, modes and cells, occupations to
  1e-12, source and interface counts equal, deposited energy to 1e-11.
- Whole runs against 16 runs of the reference (tests/golden/run_c1ls.npz) with
  the z-score rules of test_gpu_statistics, interface totals included.
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
WORKLOADS.update({"c1ls": "c1ls_source.toml"})
WEIGHTED = {"heat_source.spectrum": '"weighted"', "heat_source.branch_weights": "[1.0, 0.5, 2.0, 0.0, 0.3, 0.1]"}
GAUSS = {"heat_source.spectrum": '"gaussian"', "heat_source.frequency_center": "5.0",
         "heat_source.frequency_sigma": "1.5"}


def by_uid(c):
    live = c["mode"] >= 0
    order = np.argsort(c["uid"][live])
    return {k: v[live][order] for k, v in c.items()}


@pytest.mark.parametrize("over", [{}, WEIGHTED, GAUSS])
def test_heated_layers_step_like_the_oracle(over):
    print(f"case c1ls_source {over} n=20000", flush=True)
    p = rt.Problem(rt.config_text("c1ls_source.toml", **{"simulation.particle_count": 20000, **over}),
                   input_directory=rt.CONFIGS, mode_sampling=1)
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
        for k in ["source_updates", "step_source_updates", "interface_events", "interface_transmitted",
                  "step_absorbed", "step_injected"]:
            assert lg[k] == lo[k], k
        for k in ["step_source_energy_ev", "thermal_energy_ev", "source_pending_ev"]:
            assert lg[k] == pytest.approx(lo[k], rel=1e-11, abs=1e-9), k
    assert g.ledger()["source_updates"] > 0


def test_heated_layers_whole_runs_match_reference():
    golden = np.load(os.path.join(HERE, "golden", "run_c1ls.npz"))
    print(f"case run_c1ls n={int(golden['n'])} steps={int(golden['steps'])}", flush=True)
    temps, keff, kint, iface = gpu_runs("c1ls", golden)
    zt = z_score(temps, golden["temperature"])
    zi = z_score(iface.astype(float), golden["interface"].astype(float))
    print(f"c1ls: cell T max|z| {np.abs(zt).max():.2f}, mean z^2 {np.mean(zt ** 2):.2f}, interface z {zi}")
    assert np.abs(zt).max() <= 4.5 and np.mean(zt ** 2) <= 2.5
    assert np.all(np.abs(zi) <= 3.0)
