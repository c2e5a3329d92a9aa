"""Rough walls that are not axis-aligned on the device: the tapered fin
(taper_fin.toml, fin x walls 8.5 degrees off the vertical).

- Carrier by carrier against the oracle over 3 steps, relaxation-time walls
  (both mode-sampling modes) and strict-elastic conservative walls
  (simulation.conservative_boundaries): identical carrier sets, modes and
  cells, occupations to 1e-12, positions to 1e-6 of the extent, and the rough
  reflection counts.  The device traces the sloped triangles with its BVH and
  reflects into the table modes of the walls' own tables; the oracle carries
  explicit FP64 velocities.
- Whole runs against 16 runs of the reference (tests/golden/run_taper.npz)
  with the z-score rules of test_gpu_statistics.
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
WORKLOADS.update({"taper": "taper_fin.toml"})
STRICT = {"simulation.conservative_boundaries": "true"}


def by_uid(c):
    live = c["mode"] >= 0
    order = np.argsort(c["uid"][live])
    return {k: v[live][order] for k, v in c.items()}


@pytest.mark.parametrize("sampling,over", [(1, {}), (0, {}), (1, STRICT)])
def test_tapered_fin_steps_match_oracle(sampling, over):
    print(f"case taper_fin sampling={sampling} {over} n=20000", flush=True)
    p = rt.Problem(rt.config_text("taper_fin.toml", **{"simulation.particle_count": 20000, **over}),
                   input_directory=rt.CONFIGS, mode_sampling=sampling)
    _, _, geo, _, _, _ = p.views()
    scale = max(geo.bmax[k] - geo.bmin[k] for k in range(3))
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
        err = np.abs(a["position"].astype(np.float64) - b["position"]).max()
        assert err <= 1e-6 * scale, f"step {step}: position error {err:.3e} A"
        lg, lo = g.ledger(), o.ledger()
        for k in ["step_absorbed", "step_injected", "rough_events", "rough_specular", "rough_diffuse",
                  "live_carriers", "source_updates"]:
            assert lg[k] == lo[k], k
        for k in ["step_absorbed_energy_ev", "thermal_energy_ev", "step_source_energy_ev"]:
            assert lg[k] == pytest.approx(lo[k], rel=1e-11, abs=1e-9), k
    assert g.ledger()["rough_events"] > 0


def test_tapered_fin_whole_runs_match_reference():
    golden = np.load(os.path.join(HERE, "golden", "run_taper.npz"))
    print(f"case run_taper n={int(golden['n'])} steps={int(golden['steps'])}", flush=True)
    temps, keff, kint, _ = gpu_runs("taper", golden)
    zt = z_score(temps, golden["temperature"])
    spread = (temps.std(axis=0) > 0) | (golden["temperature"].std(axis=0) > 0)
    print(f"taper: cell T max|z| {np.abs(zt).max():.2f}, mean z^2 {np.mean(zt[spread] ** 2):.2f}")
    assert np.abs(zt).max() <= 4.5
    assert np.mean(zt[spread] ** 2) <= 2.5
