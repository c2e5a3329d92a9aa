"""HIP path against the CPU oracle, carrier by carrier, from identical Philox
populations (same counters: carrier id, step, draw, stream).

Tolerances: FP32 flight keeps positions within 1e-6 of the domain's largest
extent; hit, cell and mode indices are bit-exact; occupations and cell
temperatures agree to 1e-12 relative (FP64 tallies, summation order aside)."""
import numpy as np
import pytest

from oracle import oracle_ctypes as O
from phonomc_amd import runtime as rt

pytestmark = pytest.mark.gpu
MASK = (1 << 28) - 1


def by_uid(c):
    live = c["mode"] >= 0
    order = np.argsort(c["uid"][live])
    return {k: v[live][order] for k, v in c.items()}


def run_pair(cfg, n, steps, sampling, series=True, overrides=None):
    p = rt.Problem(rt.config_text(cfg, **{"simulation.particle_count": n, **(overrides or {})}),
                   input_directory=rt.CONFIGS, mode_sampling=sampling)
    g = rt.Context(p)
    g.set_tuning(series_on=series)
    o = O.Sim(p)
    o.initialize()
    return p, g, o


@pytest.mark.parametrize("cfg,n", [("c1_cross_plane.toml", 20000), ("c2_inplane_film.toml", 60000),
                                   ("c1c_callaway.toml", 20000), ("finfet.toml", 20000),
                                   ("c1l_layered.toml", 20000)])
def test_initial_population_identical(cfg, n):
    p, g, o = run_pair(cfg, n, 0, 1)
    a, b = by_uid(g.carriers()), o.carriers()
    assert np.array_equal(a["uid"], b["uid"])
    assert np.array_equal(a["position"].astype(np.float64), b["position"])
    assert np.array_equal(a["mode"] & MASK, b["mode"])
    assert np.array_equal(a["cell"], b["cell"])
    np.testing.assert_allclose(a["occupation"], b["occupation"], rtol=1e-14)
    np.testing.assert_allclose(g.cells()[0], o.cells()[0], rtol=0, atol=1e-10)
    assert g.ledger()["thermal_energy_ev"] == pytest.approx(o.ledger()["thermal_energy_ev"], rel=1e-13)


def test_initial_population_from_cell_field():
    """Per-cell initial field (temperature_mode 2, the warm start): the same
    population as the oracle drawn at the same field."""
    p = rt.Problem(rt.config_text("c2_inplane_film.toml", **{"simulation.particle_count": 60000}), mode_sampling=1)
    centers, _ = p.cells()
    x = centers[:, 0] / centers[:, 0].max()
    field = 300.0 + 6.0 * np.sin(np.pi * x) + 0.5 * centers[:, 2] / centers[:, 2].max()
    g = rt.Context(p, initial_field=field)
    s, m, geo, init, rough, res = p.views()
    buf = np.ascontiguousarray(field)
    init.temperature_mode = 2
    init.cell_temperature = buf.ctypes.data_as(O.C.POINTER(O.C.c_double))
    o = O.Sim(p)
    o.initialize(init)
    a, b = by_uid(g.carriers()), o.carriers()
    assert np.array_equal(a["uid"], b["uid"])
    assert np.array_equal(a["mode"] & MASK, b["mode"])
    assert np.array_equal(a["cell"], b["cell"])
    np.testing.assert_allclose(a["occupation"], b["occupation"], rtol=1e-14)
    np.testing.assert_allclose(g.cells()[0], o.cells()[0], rtol=0, atol=1e-10)


@pytest.mark.parametrize("cfg,n,sampling,series", [("c1_cross_plane.toml", 20000, 1, True),
                                                   ("c1_cross_plane.toml", 20000, 1, False),
                                                   ("c2_inplane_film.toml", 60000, 1, True),
                                                   ("c2_inplane_film.toml", 60000, 1, False),
                                                   ("c2_inplane_film.toml", 60000, 0, True),
                                                   ("c1s_source.toml", 20000, 1, True),
                                                   ("c1s_source.toml", 20000, 1, False),
                                                   ("c1c_callaway.toml", 20000, 1, True),
                                                   ("c2c_callaway.toml", 60000, 1, True),
                                                   ("c1cs_callaway_source.toml", 20000, 1, True),
                                                   ("finfet.toml", 20000, 1, True),
                                                   ("finfet.toml", 20000, 0, False),
                                                   ("c1l_layered.toml", 20000, 1, True),
                                                   ("c1l_layered.toml", 20000, 1, False),
                                                   ("finfet_callaway.toml", 20000, 1, True),
                                                   ("finfet_callaway.toml", 20000, 0, True),
                                                   ("finfet_callaway.toml", 200000, 1, True)])
def test_steps_match_oracle(cfg, n, sampling, series):
    """series=False: every pseudo-temperature evaluation is an exact carrier sum,
    the reference's own procedure; series=True: the per-cell Taylor series.
    finfet traces the 4080-triangle FinFET surface (device BVH against the
    oracle's restatement of the reference's trace), with a periodic substrate,
    a bottom reservoir, rough walls and a Gaussian heat source in the fin.
    The callaway inputs run linearized Callaway collisions (the device's Gram-space
    FP64 evaluation against the oracle's long-double restatement of the
    reference), fixed-flux reservoirs and, in c2c, strict-elastic discrete-mode
    reflections with signed occupations.  c1l_layered puts a Ge-like layer on a
    Si-like one: per-material energy lookups, diffuse-mismatch interface events
    (band search, transmit/reflect choice and occupation remap).
    finfet_callaway is the config-3 physics: Callaway N/U + isotope on the
    FinFET mesh, strict-elastic reflections on every non-reservoir facet (10 nm
    roughness), fixed-flux bottom reservoir and a 1e20 W/m^3 source; at 2e4
    carriers on the 20 x 10 x 15 grid (both sampling modes) and at 2e5 carriers
    on the full 64 x 32 x 32 grid (23 536 cells)."""
    overrides = {"simulation.grid_xyz": "[20, 10, 15]"} if cfg == "finfet_callaway.toml" and n < 100000 else None
    p, g, o = run_pair(cfg, n, 3, sampling, series, overrides)
    _, _, geo, _, _, _ = p.views()
    scale = max(geo.bmax[k] - geo.bmin[k] for k in range(3))
    worst = 0.0
    for step in range(3):
        g.step(1)
        o.step(1)
        a, b = by_uid(g.carriers()), o.carriers()
        assert np.array_equal(a["uid"], b["uid"]), f"step {step}: live carrier sets differ"
        assert np.array_equal(a["mode"] & MASK, b["mode"]), f"step {step}: modes differ"
        assert np.array_equal(a["cell"], b["cell"]), f"step {step}: cells differ"
        np.testing.assert_allclose(a["occupation"], b["occupation"], rtol=1e-12, atol=1e-300)
        err = np.abs(a["position"].astype(np.float64) - b["position"]).max()
        worst = max(worst, err / scale)
        assert err <= 1e-6 * scale, f"step {step}: position error {err:.3e} A of extent {scale:.0f} A"
        tg, eg = g.cells()
        to, eo = o.cells()
        np.testing.assert_allclose(tg, to, rtol=0, atol=1e-9)
        np.testing.assert_allclose(eg, eo, rtol=1e-12)
        lg, lo = g.ledger(), o.ledger()
        for k in ["step_absorbed", "step_injected", "total_absorbed", "total_injected", "rough_events",
                  "rough_specular", "rough_diffuse", "live_carriers", "source_updates", "interface_events",
                  "interface_transmitted", "step_source_updates"]:
            assert lg[k] == lo[k], k
        for k in ["step_absorbed_energy_ev", "step_injected_energy_ev", "thermal_energy_ev", "step_source_energy_ev",
                  "source_time_ps", "source_pending_ev"]:
            assert lg[k] == pytest.approx(lo[k], rel=1e-11, abs=1e-9), k
    led = g.ledger()
    assert (led["interface_events"] > led["interface_transmitted"] > 0) == (p.layers() is not None)
    assert (led["newton_evals"] > 0) == (p.callaway() is None)
    if not series:  # every evaluation was an exact carrier sum
        assert led["newton_passes"] >= led["newton_evals"]
    if led["source_updates"]:
        # Source deposits move a cell by far less than the series radius: the
        # series solves them all; without it every solve takes exact sums.
        if series:
            assert led["source_exact_passes"] == 0
        else:
            assert led["source_exact_passes"] > 0
    qg, qo = g.heat_flux(), o.heat_flux()
    np.testing.assert_allclose(qg, qo, rtol=1e-5, atol=1e-6 * np.abs(qo).max())
    print(f"{cfg} n={n}: largest position error {worst:.2e} of the extent")


@pytest.mark.parametrize("t_bg", [250.0, 400.0])
def test_callaway_cells_of_few_carriers_match_oracle(t_bg):
    """Config-3 physics at 2e5 carriers on the full 64 x 32 x 32 grid (about 8
    carriers per cell, many cells with dependent carrier momenta) at other
    background temperatures, with the source moved down the fin: cells whose
    normal-gain directions are (nearly) dependent take the explicit per-carrier
    Gram-Schmidt, the others the Gram-matrix form; both against the oracle's
    long-double restatement of the reference, three steps."""
    over = {"simulation.background_temperature": f"{t_bg}", "simulation.lifetime_temperature": f"{t_bg}",
            "simulation.initial_temperature": f"{t_bg}", "boundary.boundary_values": f"[{t_bg}, 10.0]",
            "heat_source.center": "[100, 50, 100]", "heat_source.sigma": "[10, 10, 10]",
            "heat_source.power_density": "1e19"}
    p, g, o = run_pair("finfet_callaway.toml", 200000, 3, 1, True, over)
    for step in range(3):
        g.step(1)
        o.step(1)
        a, b = by_uid(g.carriers()), o.carriers()
        assert np.array_equal(a["uid"], b["uid"]), f"step {step}: live carrier sets differ"
        assert np.array_equal(a["mode"] & MASK, b["mode"]), f"step {step}: modes differ"
        assert np.array_equal(a["cell"], b["cell"]), f"step {step}: cells differ"
        np.testing.assert_allclose(a["occupation"], b["occupation"], rtol=1e-12, atol=1e-300)
        tg, _ = g.cells()
        to, _ = o.cells()
        np.testing.assert_allclose(tg, to, rtol=0, atol=1e-9)
