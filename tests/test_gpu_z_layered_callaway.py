"""Linearized Callaway in layered materials on the device (c1lw_callaway.toml, with
DMM and AMM interfaces), and the nonlinear formulation (c1ln_nonlinear.toml).

- Carrier by carrier against the oracle over 3 steps: identical carrier sets,
  modes and cells, occupations to 1e-12, interface counts and thermal energy
  equal; cell temperatures through each material's C_bg to 1e-9 K.
- Whole runs against 16 runs of the reference (tests/golden/run_c1lw.npz) with
  the z-score rules of test_gpu_statistics, kappa and interface totals included.
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
WORKLOADS.update({"c1lw": "c1lw_callaway.toml"})
AMM = {"materials.interface_model": '"amm"', "materials.mass_densities_kg_m3": "[2329, 5323]",
       "materials.folders": '["synthetic:si15:callaway", "synthetic:ge15:callaway"]'}


def by_uid(c):
    live = c["mode"] >= 0
    order = np.argsort(c["uid"][live])
    return {k: v[live][order] for k, v in c.items()}


@pytest.mark.parametrize("cfg,sampling,over", [("c1lw_callaway.toml", 1, {}), ("c1lw_callaway.toml", 0, {}),
                                               ("c1lw_callaway.toml", 1, AMM), ("c1ln_nonlinear.toml", 1, {})])
def test_layered_callaway_steps_like_the_oracle(cfg, sampling, over):
    print(f"case {cfg} sampling={sampling} {over} n=20000", flush=True)
    p = rt.Problem(rt.config_text(cfg, **{"simulation.particle_count": 20000, **over}),
                   input_directory=rt.CONFIGS, mode_sampling=sampling)
    g = rt.Context(p)
    o = O.Sim(p)
    o.initialize()
    np.testing.assert_allclose(g.cells()[0], o.cells()[0], rtol=0, atol=1e-9)
    for step in range(3):
        g.step(1)
        o.step(1)
        a, b = by_uid(g.carriers()), o.carriers()
        assert np.array_equal(a["uid"], b["uid"]), f"step {step}: live carrier sets differ"
        assert np.array_equal(a["mode"] & MASK, b["mode"]), f"step {step}: modes differ"
        assert np.array_equal(a["cell"], b["cell"]), f"step {step}: cells differ"
        # the nonlinear solve (FP64 Newton on the device, long double in the oracle) agrees to 1e-10
        nonlinear = cfg.startswith("c1ln")
        np.testing.assert_allclose(a["occupation"], b["occupation"], rtol=1e-10 if nonlinear else 1e-12, atol=1e-300)
        np.testing.assert_allclose(g.cells()[0], o.cells()[0], rtol=0, atol=1e-8 if nonlinear else 1e-9)
        lg, lo = g.ledger(), o.ledger()
        for k in ["interface_events", "interface_transmitted", "step_absorbed", "step_injected"]:
            assert lg[k] == lo[k], k
        assert lg["thermal_energy_ev"] == pytest.approx(lo["thermal_energy_ev"], rel=1e-11)
    assert g.ledger()["interface_events"] > 0


def test_layered_callaway_whole_runs_match_reference():
    golden = np.load(os.path.join(HERE, "golden", "run_c1lw.npz"))
    print(f"case run_c1lw n={int(golden['n'])} steps={int(golden['steps'])}", flush=True)
    temps, keff, kint, iface = gpu_runs("c1lw", golden)
    zt = z_score(temps, golden["temperature"])
    zke, zki = z_score(keff, golden["kappa_eff"]), z_score(kint, golden["kappa_int"])
    zi = z_score(iface.astype(float), golden["interface"].astype(float))
    print(f"c1lw: cell T max|z| {np.abs(zt).max():.2f}, kappa_eff z {zke:+.2f}, kappa_int z {zki:+.2f}, "
          f"interface z {zi}")
    assert np.abs(zt).max() <= 4.5 and np.mean(zt ** 2) <= 2.5
    assert abs(zke) <= 3.0 and abs(zki) <= 3.0
    assert np.all(np.abs(zi) <= 3.0)
