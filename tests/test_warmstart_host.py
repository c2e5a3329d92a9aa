"""Host side of the warm start: the per-cell initial field of the oracle
(temperature_mode 2), the heated-film grid map and the operator features."""
import numpy as np
import pytest

from oracle import oracle_ctypes as O
from phonomc_amd import runtime as rt
from phonomc_amd import warmstart as ws


def test_oracle_cell_field_initialization():
    """A population drawn at per-cell temperatures T_c has cell temperatures T_c
    up to the sampling noise (10 000 carriers per cell: well under 0.3 K here),
    and the field differs measurably from the uniform start."""
    p = rt.Problem(rt.config_text("c1_cross_plane.toml", **{"simulation.particle_count": 200000}))
    centers, _ = p.cells()
    z = centers[:, 2] / centers[:, 2].max()
    field = 300.0 + 8.0 * np.sin(np.pi * z) - 3.0 * z
    s, m, g, init, rough, res = p.views()
    buf = np.ascontiguousarray(field)
    init.temperature_mode = 2
    init.cell_temperature = buf.ctypes.data_as(O.C.POINTER(O.C.c_double))
    o = O.Sim(p)
    o.initialize(init)
    t, _ = o.cells()
    assert np.abs(t - field).max() < 0.3
    uniform = O.Sim(p)
    uniform.initialize()
    assert np.abs(uniform.cells()[0] - field).max() > 4.0


def test_heated_film_grid_and_features():
    p = ws.sample_cases(1, 7)[0]
    problem = rt.Problem(ws.case_text(p, particles=100000))
    grid = ws.Grid(problem)
    assert sorted(grid.pos) == list(range(ws.NX * ws.NZ))
    f = ws.features(problem, p, grid)
    # The source channel carries the host's cell powers: sum density * V = total rate.
    hs = problem.heat_source()
    rate = rt.arr(hs.rate, hs.n_cells)
    total = np.sum(f[grid.pos, 0].astype(np.float64) * ws.DENSITY_SCALE * ws.EV_PER_WPS_A3 * grid.volumes)
    assert total == pytest.approx(rate.sum(), rel=1e-6)
    # Peak density equals the configured peak where a cell centre sits near the source centre.
    assert f[:, 0].max() <= p["power"] / ws.DENSITY_SCALE * (1 + 1e-6)
    assert np.all(f[:, 1] == 1.0)
    res = ws.reservoir_field(p, grid)
    np.testing.assert_allclose(f[grid.pos, 2], (res - ws.T_BG) / ws.T_SCALE, rtol=1e-6)
    # Target and initial field are inverse maps.
    t = 300.0 + np.random.default_rng(1).normal(size=len(res))
    np.testing.assert_allclose(ws.field_from_prediction(p, grid, ws.target(p, grid, t)), t, atol=1e-4)


def test_cases_are_reproducible_and_in_range():
    a, b = ws.sample_cases(50, 3), ws.sample_cases(50, 3)
    assert a == b
    for p in a:
        for k, (lo, hi) in ws.RANGES.items():
            assert lo <= p[k] <= hi
