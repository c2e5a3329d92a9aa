"""summary.txt and grid_centers.csv as the solver writes them at start-up,
against the reference's own files for the same inputs
(SimulationDomain::write_domain_summary, src/geometry/DomainOutput.cpp:12-241;
tests/golden/outputs_c1cs.npz, written by tests/golden/make_golden.py
--outputs): every section, key and value in order, the output paths aside,
and the cell centres and volumes of grid_centers.csv."""
import os
import tempfile

import numpy as np
import pytest

from phonomc_amd import runtime as rt

HERE = os.path.dirname(os.path.abspath(__file__))
WORKLOADS = {"c1": "c1_cross_plane.toml", "c2": "c2_inplane_film.toml", "c1cs": "c1cs_callaway_source.toml",
             "c1l": "c1l_layered.toml", "finfet": "finfet.toml", "finfet_cw": "finfet_callaway.toml"}
OVERRIDES = {"finfet_cw": {"simulation.grid_xyz": "[20, 10, 15]"}}
PATHS = {"output_folder", "grid_centers_csv"}


def parse(text):
    out = []
    for line in text.splitlines():
        if line.startswith("["):
            out.append((line, None))
        elif " = " in line:
            k, v = line.split(" = ", 1)
            out.append((k, None if k in PATHS else v))
    return out


@pytest.mark.parametrize("key", sorted(WORKLOADS))
def test_domain_summary_matches_the_reference(key):
    ref = np.load(os.path.join(HERE, "golden", "outputs_c1cs.npz"), allow_pickle=False)
    problem = rt.Problem(rt.config_text(WORKLOADS[key], **OVERRIDES.get(key, {})), input_directory=rt.CONFIGS)
    folder = os.path.join(tempfile.mkdtemp(), "out")
    problem.write_summary(folder)
    got = parse(open(os.path.join(folder, "summary.txt")).read())
    want = parse(str(ref["domain_summary_" + key]))
    assert [k for k, _ in got] == [k for k, _ in want]
    for (k, a), (_, b) in zip(got, want):
        assert a == b, f"{key}: {k} = {a} (reference {b})"
    g = np.loadtxt(os.path.join(folder, "grid_centers.csv"), delimiter=",", skiprows=1, ndmin=2)
    lines = str(ref["grid_centers_" + key]).splitlines()
    assert open(os.path.join(folder, "grid_centers.csv")).readline().strip() == lines[0]
    w = np.array([[float(x) for x in line.split(",")] for line in lines[1:]])
    np.testing.assert_array_equal(g[:, 0], w[:, 0])
    np.testing.assert_allclose(g[:, 1:], w[:, 1:], rtol=1e-12, atol=1e-12)
