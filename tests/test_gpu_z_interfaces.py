"""Acoustic- and mixed-mismatch interfaces of the layered Si/Ge film on the device.

- Carrier by carrier against the oracle over 3 steps from identical Philox
  populations (c1la: AMM; c1lm: MMM with half of the events on AMM): live sets,
  modes, cells and interface counters identical, occupations to 1e-12.
- Whole runs against 16 runs of the reference (tests/golden/run_c1la.npz,
  run_c1lm.npz) with the z-score rules of test_gpu_statistics, including the
  interface event and transmission totals.
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
WORKLOADS.update({"c1la": "c1la_amm.toml", "c1lm": "c1lm_mmm.toml"})


def by_uid(c):
    live = c["mode"] >= 0
    order = np.argsort(c["uid"][live])
    return {k: v[live][order] for k, v in c.items()}


@pytest.mark.parametrize("cfg", ["c1la_amm.toml", "c1lm_mmm.toml"])
def test_interface_models_step_like_the_oracle(cfg):
    print(f"case {cfg} n=20000", flush=True)
    p = rt.Problem(rt.config_text(cfg, **{"simulation.particle_count": 20000}), input_directory=rt.CONFIGS,
                   mode_sampling=1)
    g = rt.Context(p)
    o = O.Sim(p)
    o.initialize()
    _, _, geo, _, _, _ = p.views()
    scale = max(geo.bmax[k] - geo.bmin[k] for k in range(3))
    for step in range(3):
        g.step(1)
        o.step(1)
        a, b = by_uid(g.carriers()), o.carriers()
        assert np.array_equal(a["uid"], b["uid"]), f"step {step}: live carrier sets differ"
        assert np.array_equal(a["mode"] & MASK, b["mode"]), f"step {step}: modes differ"
        assert np.array_equal(a["cell"], b["cell"]), f"step {step}: cells differ"
        np.testing.assert_allclose(a["occupation"], b["occupation"], rtol=1e-12, atol=1e-300)
        assert np.abs(a["position"].astype(np.float64) - b["position"]).max() <= 1e-6 * scale
        lg, lo = g.ledger(), o.ledger()
        for k in ["step_absorbed", "step_injected", "interface_events", "interface_transmitted", "live_carriers"]:
            assert lg[k] == lo[k], k
        assert lg["thermal_energy_ev"] == pytest.approx(lo["thermal_energy_ev"], rel=1e-11)
    assert g.ledger()["interface_transmitted"] > 0


@pytest.mark.parametrize("key", ["c1la", "c1lm"])
def test_whole_run_statistics_match_reference(key):
    golden = np.load(os.path.join(HERE, "golden", f"run_{key}.npz"))
    print(f"case run_{key} n={int(golden['n'])} steps={int(golden['steps'])}", flush=True)
    temps, keff, kint, iface = gpu_runs(key, golden)
    zt = z_score(temps, golden["temperature"])
    zke = z_score(keff, golden["kappa_eff"])
    zki = z_score(kint, golden["kappa_int"])
    zi = z_score(iface.astype(float), golden["interface"].astype(float))
    print(f"{key}: cell T max|z| {np.abs(zt).max():.2f}, mean z^2 {np.mean(zt ** 2):.2f}; kappa_eff z {zke:+.2f}, "
          f"kappa_int z {zki:+.2f}; interface events z {zi[0]:+.2f}, transmitted z {zi[1]:+.2f}")
    assert abs(zke) <= 3.0 and abs(zki) <= 3.0
    assert np.all(np.abs(zi) <= 3.0)
    assert np.abs(zt).max() <= 4.5
    assert np.mean(zt ** 2) <= 2.5
