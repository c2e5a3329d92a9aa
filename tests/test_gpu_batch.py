"""Batched stepping (pmc_batch): many independent instances, one launch per kernel.

Each instance of a batch must end exactly where it would end stepped alone:
the same live carriers, modes and cells, occupations and cell temperatures to
FP64 summation order, the same ledger counts.  Checked for cross-plane films of
different thickness, background temperature and carrier count (with and
without the captured hipGraph, across the 8-step error reads and the 100-step
containment check), for FinFET meshes, for linearized Callaway films and for
instances with heat sources; each instance of a mixed batch also matches the
oracle carrier by carrier.  A failing instance latches with the reference's
message and is named.
"""
import numpy as np
import pytest

from phonomc_amd import batch as B
from phonomc_amd import runtime as rt

pytestmark = pytest.mark.gpu
MASK = (1 << 28) - 1


def by_uid(c):
    live = c["mode"] >= 0
    order = np.argsort(c["uid"][live])
    return {k: v[live][order] for k, v in c.items()}


def assert_same(a, b):
    x, y = by_uid(a.carriers()), by_uid(b.carriers())
    assert np.array_equal(x["uid"], y["uid"])
    assert np.array_equal(x["mode"], y["mode"])
    assert np.array_equal(x["cell"], y["cell"])
    assert np.array_equal(x["position"], y["position"])
    np.testing.assert_allclose(x["occupation"], y["occupation"], rtol=1e-12, atol=1e-300)
    ta, _ = a.cells()
    tb, _ = b.cells()
    np.testing.assert_allclose(ta, tb, rtol=0, atol=1e-9)
    la, lb = a.ledger(), b.ledger()
    for k in ["step", "live_carriers", "total_absorbed", "total_injected", "rough_events", "rough_specular",
              "rough_diffuse", "carrier_steps"]:
        assert la[k] == lb[k], k
    for k in ["thermal_energy_ev", "total_absorbed_energy_ev", "total_injected_energy_ev"]:
        assert la[k] == pytest.approx(lb[k], rel=1e-11, abs=1e-9), k


@pytest.mark.parametrize("graph", [True, False])
def test_batch_equals_independent_contexts(graph):
    inst = B.film_ensemble(6, seed=3, carriers_at_10nm=2000, thickness_nm=(10.0, 200.0))
    pa, ca = B.contexts(inst)
    pb, cb = B.contexts(inst)
    batch = rt.Batch(ca, graph=graph)
    batch.step(25)
    for c in cb:
        c.step(25)
    for a, b in zip(ca, cb):
        assert_same(a, b)
    # The instances differ: different temperatures and carrier counts.
    assert len({a.ledger()["live_carriers"] for a in ca}) == len(ca)
    batch.close()


def test_batch_crosses_the_containment_check():
    inst = B.film_ensemble(3, seed=4, carriers_at_10nm=1000, thickness_nm=(20.0, 60.0))
    _, ca = B.contexts(inst)
    _, cb = B.contexts(inst)
    batch = rt.Batch(ca)
    batch.step(105)
    for c in cb:
        c.step(105)
    for a, b in zip(ca, cb):
        assert_same(a, b)


def test_batch_of_meshes_and_of_callaway_films():
    text = rt.config_text("finfet.toml", **{"simulation.particle_count": 100000, "heat_source.enabled": "false"})
    ps = [rt.Problem(text.replace("random_seed = 12345", f"random_seed = {s}"), input_directory=rt.CONFIGS)
          for s in (1, 2)]
    ca = [rt.Context(p) for p in ps]
    cb = [rt.Context(p) for p in ps]
    rt.Batch(ca).step(6)
    for c in cb:
        c.step(6)
    for a, b in zip(ca, cb):
        assert_same(a, b)
    ps = [rt.Problem(rt.config_text("c1c_callaway.toml", **{"simulation.particle_count": n,
                                                             "simulation.random_seed": s}))
          for n, s in ((20000, 5), (30000, 6))]
    ca = [rt.Context(p) for p in ps]
    cb = [rt.Context(p) for p in ps]
    rt.Batch(ca).step(6)
    for c in cb:
        c.step(6)
    for a, b in zip(ca, cb):
        assert_same(a, b)


def test_batch_instances_match_the_oracle():
    """Config 4 against the oracle: each instance of a mixed batch (film
    thickness, background temperature, carrier count and seed differ; alias and
    inverse-CDF reservoir sampling) carrier by carrier against an oracle run of
    the same input over three steps: the same live carriers, modes and cells,
    occupations to 1e-12, cell temperatures to 1e-9 K, the same ledger counts."""
    from oracle import oracle_ctypes as O

    inst = B.film_ensemble(4, seed=11, carriers_at_10nm=1500, thickness_nm=(10.0, 120.0))
    problems = [rt.Problem(i["text"], mode_sampling=k % 2) for k, i in enumerate(inst)]
    ctx = [rt.Context(p) for p in problems]
    orc = [O.Sim(p) for p in problems]
    for o in orc:
        o.initialize()
    batch = rt.Batch(ctx)
    for step in range(3):
        batch.step(1)
        for k, (c, o) in enumerate(zip(ctx, orc)):
            o.step(1)
            a, b = by_uid(c.carriers()), o.carriers()
            assert np.array_equal(a["uid"], b["uid"]), (k, step)
            assert np.array_equal(a["mode"] & MASK, b["mode"]), (k, step)
            assert np.array_equal(a["cell"], b["cell"]), (k, step)
            np.testing.assert_allclose(a["occupation"], b["occupation"], rtol=1e-12, atol=1e-300)
            np.testing.assert_allclose(c.cells()[0], o.cells()[0], rtol=0, atol=1e-9)
            lg, lo = c.ledger(), o.ledger()
            for key in ["live_carriers", "total_absorbed", "total_injected", "step_absorbed", "step_injected"]:
                assert lg[key] == lo[key], (k, step, key)
    batch.close()


def test_batch_with_heat_sources_equals_independent_contexts():
    """Instances with prescribed heat sources (device-side half-step time
    factors, so the captured graph replays every step): a cross-plane film with
    a Gaussian source, the same film with twice the power, one without a source,
    and two FinFET Callaway instances with different source positions; each
    ends where it ends stepped alone, and deposits its own power."""
    texts = [rt.config_text("c1s_source.toml", **{"simulation.particle_count": 20000}),
             rt.config_text("c1s_source.toml", **{"simulation.particle_count": 15000,
                                                  "heat_source.power_density": "6e17",
                                                  "simulation.random_seed": 77}),
             rt.config_text("c1_cross_plane.toml", **{"simulation.particle_count": 10000})]
    ca = [rt.Context(rt.Problem(t)) for t in texts]
    cb = [rt.Context(rt.Problem(t)) for t in texts]
    batch = rt.Batch(ca)
    batch.step(12)
    for c in cb:
        c.step(12)
    for a, b in zip(ca, cb):
        assert_same(a, b)
        la, lb = a.ledger(), b.ledger()
        assert la["total_source_energy_ev"] == pytest.approx(lb["total_source_energy_ev"], rel=1e-12)
        assert abs(la["step_balance_residual_ev"]) <= 1e-12 * la["thermal_energy_ev"]
    assert ca[0].ledger()["total_source_energy_ev"] > 0 and ca[2].ledger()["total_source_energy_ev"] == 0
    assert ca[1].ledger()["total_source_energy_ev"] > ca[0].ledger()["total_source_energy_ev"]
    batch.close()
    fin = [rt.config_text("finfet_callaway.toml", **{"simulation.particle_count": 60000,
                                                     "simulation.grid_xyz": "[20, 10, 15]",
                                                     "heat_source.center": f"[100, 50, {z}]",
                                                     "simulation.random_seed": 100 + z}) for z in (90, 140)]
    ca = [rt.Context(rt.Problem(t, input_directory=rt.CONFIGS)) for t in fin]
    cb = [rt.Context(rt.Problem(t, input_directory=rt.CONFIGS)) for t in fin]
    batch = rt.Batch(ca)
    batch.step(8)
    for c in cb:
        c.step(8)
    for a, b in zip(ca, cb):
        assert_same(a, b)
    batch.close()


def test_config4_ensemble_of_512_films():
    """Config 4 at its size: 512 cross-plane films (thickness log-uniform
    10-1000 nm, T_bg 150-450 K, +-1 K reservoirs, carriers proportional to
    thickness, 1.1e8 in all) stepped as one batch.  Every instance closes its
    energy balance every step (1e-12 of its thermal energy) and keeps its carrier
    count within a few percent; eight sampled instances stepped alone from the
    same inputs reach the same kappa_eff as in the batch (to FP64 summation
    order, so well inside 3 standard errors)."""
    inst = B.film_ensemble(512, seed=7, carriers_at_10nm=10000)
    problems, ctx = B.contexts(inst)
    assert sum(c.ledger()["live_carriers"] for c in ctx) > 1.0e8
    batch = rt.Batch(ctx)
    for _ in range(40):
        batch.step(1)
        for c, i in zip(ctx, inst):
            led = c.ledger()
            assert abs(led["step_balance_residual_ev"]) <= 1e-12 * led["thermal_energy_ev"]
    for c, i in zip(ctx, inst):
        assert abs(c.ledger()["live_carriers"] - i["carriers"]) < 0.05 * i["carriers"] + 50
    batch.close()
    rng = np.random.default_rng(1)
    for k in rng.choice(512, 8, replace=False):
        p = rt.Problem(inst[k]["text"])
        alone = rt.Context(p)
        alone.step(40)
        ka, kb = B.kappa_eff(problems[k], ctx[k]), B.kappa_eff(p, alone)
        assert ka == pytest.approx(kb, rel=1e-9, abs=1e-9), k


def test_batch_owns_its_contexts():
    """While a batch exists its contexts refuse to be stepped, tuned or
    reconfigured alone; after pmc_batch_destroy they step again."""
    ctx = [rt.Context(rt.Problem(rt.config_text("c1_cross_plane.toml", **{"simulation.particle_count": 3000})))
           for _ in range(2)]
    batch = rt.Batch(ctx)
    with pytest.raises(RuntimeError, match="belongs to a batch"):
        ctx[0].step(1)
    with pytest.raises(RuntimeError, match="belongs to a batch"):
        ctx[1].set_launch_config(2, 512)
    with pytest.raises(RuntimeError, match="belongs to a batch"):
        ctx[0].set_tuning(series_on=False)
    with pytest.raises(RuntimeError):
        rt.Batch([ctx[0]])
    batch.step(2)
    batch.close()
    ctx[0].step(1)
    assert ctx[0].ledger()["step"] == 3


def test_batch_failure_names_the_instance_and_latches():
    good = rt.config_text("c1_cross_plane.toml", **{"simulation.particle_count": 5000})
    bad = rt.config_text("c1_cross_plane.toml", **{"simulation.particle_count": 2000, "simulation.time_step": "50.0",
                                                   "geometry.sizes": "[1, 1, 1]", "simulation.grid_xyz": "[1, 1, 2]"})
    ctx = [rt.Context(rt.Problem(good)), rt.Context(rt.Problem(bad))]
    batch = rt.Batch(ctx)
    with pytest.raises(RuntimeError, match="instance 1: Particle .* exceeded 64 transport events"):
        batch.step(1)
    with pytest.raises(RuntimeError, match="pmc error 8"):
        batch.step(1)


def test_batch_rejects_mixed_instances():
    film = rt.Context(rt.Problem(rt.config_text("c1_cross_plane.toml", **{"simulation.particle_count": 5000})))
    cw = rt.Context(rt.Problem(rt.config_text("c1c_callaway.toml", **{"simulation.particle_count": 5000})))
    with pytest.raises(RuntimeError, match="collision model"):
        rt.Batch([film, cw])
    other = rt.Context(rt.Problem(rt.config_text("c1_cross_plane.toml", **{"simulation.particle_count": 5000})))
    other.step(1)
    with pytest.raises(RuntimeError, match="different steps"):
        rt.Batch([film, other])


@pytest.mark.parametrize("cfg,n", [("c2c_callaway.toml", 100000), ("finfet.toml", 200000),
                                   ("c2_inplane_film.toml", 200000)])
def test_autotuned_launch_changes_no_result(cfg, n):
    """pmc_autotune times each candidate on real steps and keeps the fastest; a
    context tuned that way ends where an untuned one with the same seed ends after
    the same number of steps."""
    text = rt.config_text(cfg, **{"simulation.particle_count": n})
    p = rt.Problem(text, input_directory=rt.CONFIGS)
    a, b = rt.Context(p), rt.Context(p)
    chosen, report = a.autotune(steps=2)
    assert chosen["event_split"] in (1, 2, 4, 8)
    assert chosen["callaway_threads"] in (256, 384, 1024)
    assert len([k for k in report if k.startswith("event_split")]) == 4
    if p.callaway() is not None:
        assert "callaway_threads=1024" in report
    steps = a.ledger()["step"]
    assert steps >= 8
    b.step(steps)
    assert_same(a, b)
    print(cfg, chosen, {k: round(v, 4) for k, v in report.items()})
    # The configuration can also be set by hand.
    b.set_launch_config(2, 384)
    assert b.launch_config() == {"event_split": 2, "callaway_threads": 384}
    with pytest.raises(RuntimeError, match="event_split"):
        b.set_launch_config(3, 512)


def test_batch_of_many_finfets_splits_its_launches():
    """96 heated FinFET instances (linearized Callaway, dt 1 ps, 2e4 carriers):
    the heat-source kernels take one workgroup per carrier chunk, about 2.4e4
    per instance, so their batched grid exceeds one launch's 2^21 workgroups
    and is split; sampled instances still end where they end stepped alone."""
    from phonomc_amd import warmstart3d as w3
    cases = w3.sample_cases(96, 17)
    problems = [w3.problem_of(p, 20000, 300 + k) for k, p in enumerate(cases)]
    ca = [rt.Context(p) for p in problems]
    batch = rt.Batch(ca)
    batch.step(6)
    for k in (0, 47, 95):
        alone = rt.Context(problems[k])
        alone.step(6)
        assert_same(ca[k], alone)
        alone.close()
    batch.close()
