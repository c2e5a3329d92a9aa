"""Output files of the drop-in solver against the reference's own for the same
input (tests/golden/outputs_c1cs.npz: one reference run of the heated Callaway
film, 2e4 carriers, 200 steps, rows every 50, written by
tests/golden/make_golden.py --outputs).

The random streams differ (Philox here, mt19937_64 in the reference), so
values that depend on them are compared by key only; what the input fixes is
compared by value: the convergence.txt header and its step / time columns, the
prescribed heat-source energy of every row (the source's total rate times the
elapsed on-time), the pending energy and the ledger residual; every line of
heat_source_ledger.txt but the carrier-dependent deposit split; and the summary
sections [escaped_particle_recovery], [reservoir_balance_diagnostics],
[energy_weighting] and [local_heat_source_occupation_injection]
(src/solver/ResultWriter.cpp:100-170, :290-354)."""
import os
import tempfile

import numpy as np
import pytest

from phonomc_amd import runtime as rt

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
SECTIONS = ["escaped_particle_recovery", "reservoir_balance_diagnostics", "energy_weighting",
            "local_heat_source_occupation_injection"]
# keys whose values depend on the random streams
STOCHASTIC = {"absorbed_particles_total", "injected_particles_total", "net_particles_total",
              "absorbed_deviational_energy_total_ev", "injected_deviational_energy_total_ev",
              "current_particle_count", "total_thermal_energy_ev", "lifetime_energy_residual_total_ev",
              "energy_balance_residual_total_ev", "occupation_updates_total", "occupation_updates_last_step",
              "injected_energy_total_ev", "injected_energy_last_step_ev", "source_ledger_residual_ev"}


def sections(text):
    out, cur = {}, None
    for line in text.splitlines():
        line = line.strip()
        if line.startswith("[") and line.endswith("]"):
            cur = line[1:-1]
            out[cur] = []
        elif cur is not None and " = " in line:
            k, v = line.split(" = ", 1)
            out[cur].append((k, v))
    return out


def same_value(a, b):
    try:
        x, y = float(a), float(b)
    except ValueError:
        return a == b
    return x == pytest.approx(y, rel=1e-12, abs=1e-12)


@pytest.fixture(scope="module")
def runs():
    ref = np.load(os.path.join(HERE, "golden", "outputs_c1cs.npz"), allow_pickle=False)
    tmp = tempfile.mkdtemp()
    folder = os.path.join(tmp, "out")
    text = rt.config_text("c1cs_callaway_source.toml", **{
        "simulation.particle_count": int(ref["n"]), "simulation.iterations": int(ref["steps"]),
        "simulation.random_seed": int(ref["seed"]), "simulation.convergence_write_interval": int(ref["interval"]),
        "io.output_folder": f'"{folder}"'})
    solver = rt.Solver(rt.Problem(text, input_directory=rt.CONFIGS), output_folder=folder)
    solver.run(int(ref["steps"]))
    summary = os.path.join(tmp, "profile.txt")
    solver.append_summary(summary)
    solver.close()
    return ref, folder, open(summary).read()


def test_convergence_columns_match_the_reference(runs):
    ref, folder, _ = runs
    lines = open(os.path.join(folder, "convergence.txt")).read().splitlines()
    assert lines[0] == str(ref["convergence_header"])
    rows = np.array([[float(x) for x in line.split()] for line in lines[1:]])
    want = ref["convergence"]
    assert rows.shape == want.shape
    np.testing.assert_array_equal(rows[:, 0], want[:, 0])        # timestep
    np.testing.assert_allclose(rows[:, 1], want[:, 1], rtol=1e-12)  # time_ps
    names = lines[0].lstrip("# ").split()
    col = {n: i for i, n in enumerate(names)}
    np.testing.assert_allclose(rows[:, col["hs_prescribed_total_ev"]], want[:, col["hs_prescribed_total_ev"]],
                               rtol=1e-12)
    np.testing.assert_allclose(rows[:, col["hs_pending_total_ev"]], want[:, col["hs_pending_total_ev"]], atol=1e-9)
    prescribed = rows[:, col["hs_prescribed_total_ev"]]
    assert np.all(np.abs(rows[:, col["hs_ledger_residual_ev"]]) <= 1e-9 * np.maximum(prescribed, 1.0))
    # per-step source work of the same order as the reference's (different carriers)
    upd, e = col["hs_occupation_updates"], col["hs_injected_energy_ev"]
    assert np.all(rows[1:, upd] > 0) and np.all(rows[1:, e] > 0)
    np.testing.assert_allclose(rows[1:, e], want[1:, e], rtol=1e-9)   # the step's deposit is the step's accrual
    np.testing.assert_allclose(rows[1:, upd], want[1:, upd], rtol=0.2)


def test_heat_source_ledger_matches_the_reference(runs):
    ref, folder, _ = runs
    got = open(os.path.join(folder, "heat_source_ledger.txt")).read().splitlines()
    want = str(ref["source_ledger"]).splitlines()
    assert got[:3] == want[:3] and len(got) == len(want)
    g = np.array([[float(x) for x in line.split()] for line in got[3:]])
    w = np.array([[float(x) for x in line.split()] for line in want[3:]])
    np.testing.assert_array_equal(g[:, 0], w[:, 0])                     # cell
    np.testing.assert_allclose(g[:, 1:4], w[:, 1:4], rtol=1e-12)        # centre
    np.testing.assert_array_equal(g[:, 4], w[:, 4])                     # material
    np.testing.assert_allclose(g[:, 5:9], w[:, 5:9], rtol=1e-12, atol=1e-9)  # power, prescribed, deposited, pending


def test_summary_sections_match_the_reference(runs):
    ref, _, text = runs
    got, want = sections(text), sections(str(ref["summary"]))
    for name in SECTIONS:
        assert name in got, name
        assert [k for k, _ in got[name]] == [k for k, _ in want[name]], name
        for (k, a), (_, b) in zip(got[name], want[name]):
            if k not in STOCHASTIC:
                assert same_value(a, b), f"[{name}] {k}: {a} vs {b}"
    order = [s for s in got if s in SECTIONS]
    assert order == [s for s in want if s in SECTIONS]
