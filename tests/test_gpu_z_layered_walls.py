"""Rough side walls of a layered box on the device (c1lr_rough.toml, and
c1lc_conservative.toml with strict-elastic conservative walls): each material
reflects through its own table (pmc_layers.facet_rough).

- Carrier by carrier against the oracle over 3 steps, both mode-sampling modes:
  identical carrier sets, modes and cells, occupations to 1e-12, rough and
  interface counts equal.
- Whole runs against 16 runs of the reference (tests/golden/run_c1lr.npz) with
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
WORKLOADS.update({"c1lr": "c1lr_rough.toml", "c1lc": "c1lc_conservative.toml"})


def by_uid(c):
    live = c["mode"] >= 0
    order = np.argsort(c["uid"][live])
    return {k: v[live][order] for k, v in c.items()}


@pytest.mark.parametrize("cfg,sampling", [("c1lr_rough.toml", 1), ("c1lr_rough.toml", 0),
                                          ("c1lc_conservative.toml", 1)])
def test_layered_rough_walls_step_like_the_oracle(cfg, sampling):
    print(f"case {cfg} sampling={sampling} n=20000", flush=True)
    p = rt.Problem(rt.config_text(cfg, **{"simulation.particle_count": 20000}),
                   input_directory=rt.CONFIGS, mode_sampling=sampling)
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
        for k in ["step_absorbed", "step_injected", "rough_events", "rough_specular", "rough_diffuse",
                  "interface_events", "interface_transmitted", "live_carriers"]:
            assert lg[k] == lo[k], k
        assert lg["thermal_energy_ev"] == pytest.approx(lo["thermal_energy_ev"], rel=1e-11)
    assert g.ledger()["rough_events"] > 0 and g.ledger()["interface_events"] > 0


@pytest.mark.parametrize("key", ["c1lr", "c1lc"])
def test_layered_rough_walls_whole_runs_match_reference(key):
    golden = np.load(os.path.join(HERE, "golden", f"run_{key}.npz"))
    print(f"case run_{key} n={int(golden['n'])} steps={int(golden['steps'])}", flush=True)
    temps, keff, kint, iface = gpu_runs(key, golden)
    zt = z_score(temps, golden["temperature"])
    zke, zki = z_score(keff, golden["kappa_eff"]), z_score(kint, golden["kappa_int"])
    zi = z_score(iface.astype(float), golden["interface"].astype(float))
    print(f"{key}: cell T max|z| {np.abs(zt).max():.2f}, kappa_eff z {zke:+.2f}, kappa_int z {zki:+.2f}, "
          f"interface z {zi}")
    assert np.abs(zt).max() <= 4.5 and np.mean(zt ** 2) <= 2.5
    assert abs(zke) <= 3.0 and abs(zki) <= 3.0
    assert np.all(np.abs(zi) <= 3.0)
