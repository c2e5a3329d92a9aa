"""Nonlinear Callaway scattering and conservative RTA walls on the device.

- Carrier by carrier against the oracle over 3 steps from identical Philox
  populations: c1cn (nonlinear Callaway, cross-plane film, fixed rate
  temperature), c2cn (nonlinear Callaway, in-plane film with strict-elastic
  rough walls, rates at the cell temperature) and c2r (RTA with
  simulation.conservative_boundaries).  Indices and cells are bit-exact;
  occupations agree to 1e-10 relative for nonlinear Callaway (the device's FP64
  moment solve against the oracle's long-double one, both converged to the
  reference's 2e-13 gradient tolerance) and 1e-12 for RTA.
- Whole runs against 16 runs of the reference (tests/golden/run_c1cn.npz,
  run_c2cn.npz, run_c2r.npz), with the z-score rules of test_gpu_statistics.
- The nonlinear film at 1e6 carriers: energy closure every step and
  non-negative occupations.
"""
import os

import numpy as np
import pytest

from oracle import oracle_ctypes as O
from phonomc_amd import runtime as rt

from test_gpu_statistics import gpu_runs, z_score, WORKLOADS

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
MASK = (1 << 28) - 1
WORKLOADS.update({"c1cn": "c1cn_nonlinear.toml", "c2cn": "c2cn_nonlinear.toml", "c2r": "c2r_conservative.toml"})


def by_uid(c):
    live = c["mode"] >= 0
    order = np.argsort(c["uid"][live])
    return {k: v[live][order] for k, v in c.items()}


@pytest.mark.parametrize("cfg,n,rtol", [("c1cn_nonlinear.toml", 20000, 1e-10),
                                        ("c2cn_nonlinear.toml", 60000, 1e-10),
                                        ("c2r_conservative.toml", 60000, 1e-12)])
def test_nonlinear_and_conservative_steps_match_oracle(cfg, n, rtol):
    print(f"case {cfg} n={n}", flush=True)
    p = rt.Problem(rt.config_text(cfg, **{"simulation.particle_count": n}), input_directory=rt.CONFIGS,
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
        np.testing.assert_allclose(a["occupation"], b["occupation"], rtol=rtol, atol=1e-300)
        assert np.abs(a["position"].astype(np.float64) - b["position"]).max() <= 1e-6 * scale
        tg, eg = g.cells()
        to, eo = o.cells()
        np.testing.assert_allclose(tg, to, rtol=0, atol=1e-8)
        lg, lo = g.ledger(), o.ledger()
        for k in ["step_absorbed", "step_injected", "total_absorbed", "total_injected", "rough_events",
                  "rough_specular", "rough_diffuse", "live_carriers"]:
            assert lg[k] == lo[k], k
        for k in ["step_absorbed_energy_ev", "step_injected_energy_ev", "thermal_energy_ev"]:
            assert lg[k] == pytest.approx(lo[k], rel=1e-10, abs=1e-8), k
        assert abs(lg["step_balance_residual_ev"]) <= 1e-12 * lg["thermal_energy_ev"]


def test_nonlinear_film_closes_energy_at_1e6_carriers():
    print("case c2cn_nonlinear.toml n=1000000 steps=40", flush=True)
    p = rt.Problem(rt.config_text("c2cn_nonlinear.toml"), input_directory=rt.CONFIGS)
    g = rt.Context(p)
    for _ in range(40):
        g.step(1)
        L = g.ledger()
        assert abs(L["step_balance_residual_ev"]) <= 1e-12 * L["thermal_energy_ev"]
    c = g.carriers()
    live = c["mode"] >= 0
    assert np.all(c["occupation"][live] >= 0.0)
    assert abs(L["live_carriers"] - 1_000_000) < 0.05 * 1_000_000


@pytest.mark.parametrize("key", ["c1cn", "c2cn", "c2r"])
def test_whole_run_statistics_match_reference(key):
    path = os.path.join(HERE, "golden", f"run_{key}.npz")
    golden = np.load(path)
    print(f"case run_{key} n={int(golden['n'])} steps={int(golden['steps'])}", flush=True)
    temps, keff, kint, _ = gpu_runs(key, golden)
    zt = z_score(temps, golden["temperature"])
    zke = z_score(keff, golden["kappa_eff"])
    zki = z_score(kint, golden["kappa_int"])
    print(f"{key}: cell T max|z| {np.abs(zt).max():.2f}, mean z^2 {np.mean(zt ** 2):.2f}; kappa_eff "
          f"{keff.mean():.3f} vs {golden['kappa_eff'].mean():.3f} (z {zke:+.2f}); kappa_int {kint.mean():.3f} vs "
          f"{golden['kappa_int'].mean():.3f} (z {zki:+.2f})")
    assert abs(zke) <= 3.0 and abs(zki) <= 3.0
    assert np.abs(zt).max() <= 4.5
    assert np.mean(zt ** 2) <= 2.5
