"""Batched stepping (pmc_batch): many independent instances, one launch per kernel.

Each instance of a batch must end exactly where it would end stepped alone:
the same live carriers, modes and cells, occupations and cell temperatures to
FP64 summation order, the same ledger counts.  Checked for cross-plane films of
different thickness, background temperature and carrier count (with and
without the captured hipGraph, across the 8-step error reads and the 100-step
containment check), for FinFET meshes and for linearized Callaway films.  A
failing instance latches with the reference's message and is named.
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


def test_batch_rejects_mixed_or_sourced_instances():
    film = rt.Context(rt.Problem(rt.config_text("c1_cross_plane.toml", **{"simulation.particle_count": 5000})))
    src = rt.Context(rt.Problem(rt.config_text("c1s_source.toml", **{"simulation.particle_count": 5000})))
    with pytest.raises(RuntimeError, match="heat source"):
        rt.Batch([film, src])
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
    assert chosen["callaway_threads"] in (128, 256, 512)
    assert len([k for k in report if k.startswith("event_split")]) == 4
    if p.callaway() is not None:
        assert "callaway_threads=512" in report
    steps = a.ledger()["step"]
    assert steps >= 8
    b.step(steps)
    assert_same(a, b)
    print(cfg, chosen, {k: round(v, 4) for k, v in report.items()})
    # The configuration can also be set by hand.
    b.set_launch_config(2, 512)
    assert b.launch_config() == {"event_split": 2, "callaway_threads": 512}
    with pytest.raises(RuntimeError, match="event_split"):
        b.set_launch_config(3, 512)
