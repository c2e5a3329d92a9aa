"""The CPU oracle against known-answer vectors of the reference's own functions
(tests/golden/*.npz, written by tests/golden/make_golden.py) and published
Philox4x32-10 test vectors."""
import numpy as np
import pytest

from oracle import oracle_ctypes as O

KAT = [  # Random123 kat_vectors, philox4x32 R=10
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,expected", KAT)
def test_philox_known_answers(ctr, key, expected):
    assert O.philox(ctr, key) == expected


def test_u53_range_and_resolution():
    lib = O.lib()
    assert lib.orc_u53(0, 0) == 0.0
    assert lib.orc_u53(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0 ** -53
    assert lib.orc_u53(1 << 11, 0) == 2.0 ** -53


@pytest.mark.parametrize("key", ["c1", "c2"])
def test_trace_box_and_grid_index_bit_exact(problems, golden, key):
    p, g = problems(key), golden(f"domain_{key}.npz")
    for k in range(len(g["trace_pos"])):
        f, t, pt = O.trace_box(p, g["trace_pos"][k], g["trace_vel"][k])
        assert f == g["trace_facet"][k]
        assert t == g["trace_t"][k]
        assert np.array_equal(pt, g["trace_point"][k])
    cells = [O.grid_index(p, x) for x in g["grid_pos"]]
    assert np.array_equal(cells, g["grid_cell"])


def test_rough_sampling_bit_exact(problems, golden):
    p, g = problems("c2"), golden("rough_c2.npz")
    for facet, table in [(4, 0), (5, 1)]:
        ins, occ, u = g[f"f{facet}_in"], g[f"f{facet}_occ"], g[f"f{facet}_u"]
        out, out_occ = g[f"f{facet}_out"], g[f"f{facet}_out_occ"]
        for k in range(len(ins)):
            a, o, kind, used = O.rough_sample(p, table, ins[k], occ[k], [u[k, 0], u[k, 1], 0.5, 0.5])
            assert a == out[k]
            assert o == pytest.approx(out_occ[k], rel=1e-12, abs=0)


@pytest.mark.parametrize("key,dt", [("c1", 1.0), ("c2", 0.5)])
def test_reservoir_draws_bit_exact(problems, golden, key, dt):
    p, g = problems(key), golden(f"reservoir_{key}.npz")
    s, m, geo, i, rough, res = p.views()
    for r in range(len(res)):
        assert res[r].initial_count == int(g[f"r{r}_initial"])
        u = g[f"r{r}_u"]
        modes = [O.reservoir_mode(p, r, u[k, 0], 1) for k in range(len(u))]
        assert np.array_equal(modes, g[f"r{r}_mode"])
        assert np.array_equal(dt * u[:, 1], g[f"r{r}_time"])
        su = g[f"r{r}_surf_u"]
        pts = np.array([O.surface_point(p, r, *su[k]) for k in range(len(su))])
        assert np.array_equal(pts, g[f"r{r}_surf_point"])


@pytest.mark.parametrize("key", ["c1", "c2"])
def test_refresh_and_rta_match_reference(problems, golden, key):
    p, g = problems(key), golden(f"collision_{key}.npz")
    e, t = O.refresh(p, g["active"], g["cell"], g["occ"])
    np.testing.assert_allclose(e, g["cell_energy"], rtol=1e-14, atol=0)
    np.testing.assert_allclose(t, g["cell_t"], rtol=0, atol=1e-11)
    occ, res, pseudo = O.rta(p, g["cell_t"], g["active"], g["cell"], g["occ"], float(g["dt"]), int(g["local"]),
                             float(g["lifetime_t"]))
    np.testing.assert_allclose(occ, g["occ_after"], rtol=1e-12, atol=1e-300)
    assert res == pytest.approx(float(g["residual"]), rel=1e-9, abs=1e-12)


NARROW = {"heat_source.frequency_center": "40.0", "heat_source.frequency_sigma": "0.05"}


@pytest.mark.parametrize("cfg,fixture,over", [("c1s_source.toml", "source_c1s.npz", {}),
                                              ("c1cs_callaway_source.toml", "source_c1cs.npz", {}),
                                              ("c1sw_weighted.toml", "source_c1sw.npz", {}),
                                              ("c1sg_gaussian.toml", "source_c1sg.npz", {}),
                                              ("c1sg_gaussian.toml", "source_c1sg_narrow.npz", NARROW)])
def test_source_deposit_matches_reference(golden, cfg, fixture, over):
    """Heat-source deposit against PhononSourceOperator::apply: the thermal
    spectrum under RTA (a bracket + Newton pseudo-temperature per cell) and under
    linearized Callaway (increments proportional to the mode heat capacity at
    T_bg), the weighted spectrum (branch weights over the largest, branch 3 off)
    and the Gaussian spectrum in frequency (5 THz centre, 1.5 THz width; and a
    0.05 THz Gaussian at 40 THz, above every mode, whose weights survive only by
    the per-cell peak normalization); all with the roundoff correction on the
    largest carrier."""
    from phonomc_amd import runtime as rt

    g = golden(fixture)
    p = rt.Problem(rt.config_text(cfg, **over), input_directory=rt.CONFIGS, mode_sampling=1)
    pending = g["rates"] * float(g["factor"])
    occ, energy, updates, left = O.source_deposit(p, pending, g["cell_t"], g["active"], g["cell"], g["occ"])
    np.testing.assert_allclose(occ, g["occ_after"], rtol=1e-12, atol=0)
    assert energy == pytest.approx(float(g["deposited"]), rel=1e-14)
    assert updates == int(g["updates"])
    assert np.all(left == 0.0)
    # the deposit is exact: sum of hw * increment * weight equals the budget
    _, m, _, _, _, _ = p.views()
    omega = np.ctypeslib.as_array(m.omega, shape=(m.n_active,))
    dep = np.sum(6.582119569e-4 * omega[g["active"]] * (occ - g["occ"])) * m.energy_weight
    assert dep == pytest.approx(float(g["deposited"]), rel=1e-11)
    cw = p.callaway()
    if cw is not None:
        # Linearized: the energy increment of each carrier is proportional to its
        # mode heat capacity, one factor per cell (all carriers but the corrected one).
        cap = rt.arr(cw.capacity, m.n_active)[g["active"]]
        ratio = 6.582119569e-4 * omega[g["active"]] * (occ - g["occ"]) / cap
        for c in np.unique(g["cell"]):
            r = ratio[g["cell"] == c]
            if np.all(r == 0.0):
                continue
            med = np.median(r)
            assert np.sum(~np.isclose(r, med, rtol=1e-9, atol=0)) <= 1


def test_planted_mistake_is_caught(problems, golden):
    """A plausible slip (window sampled from the wrong uniform) fails the rough vectors."""
    p, g = problems("c2"), golden("rough_c2.npz")
    ins, occ, u, out = g["f4_in"], g["f4_occ"], g["f4_u"], g["f4_out"]
    wrong = 0
    for k in range(400):
        a, _, _, _ = O.rough_sample(p, 0, ins[k], occ[k], [u[k, 0], 1.0 - u[k, 1], 0.5, 0.5])
        wrong += a != out[k]
    assert wrong > 0


def test_source_spectrum_planted_mistake_is_caught(golden):
    """A Gaussian centre read in rad/ps instead of THz (2 pi times too high)
    moves the deposit far from the reference's; so does depositing the weighted
    budget thermally."""
    from phonomc_amd import runtime as rt

    g = golden("source_c1sg.npz")
    wrong = rt.Problem(rt.config_text("c1sg_gaussian.toml", **{"heat_source.frequency_center": str(5.0 * 2 * np.pi)}),
                       input_directory=rt.CONFIGS, mode_sampling=1)
    occ, *_ = O.source_deposit(wrong, g["rates"] * float(g["factor"]), g["cell_t"], g["active"], g["cell"], g["occ"])
    assert np.max(np.abs(occ - g["occ_after"]) / np.abs(g["occ_after"])) > 1e-3
    g = golden("source_c1sw.npz")
    thermal = rt.Problem(rt.config_text("c1sw_weighted.toml").replace('spectrum = "weighted"', 'spectrum = "thermal"')
                         .replace("branch_weights = [1.0, 0.5, 2.0, 0.0, 0.3, 0.1]\n", ""),
                         input_directory=rt.CONFIGS, mode_sampling=1)
    occ, *_ = O.source_deposit(thermal, g["rates"] * float(g["factor"]), g["cell_t"], g["active"], g["cell"], g["occ"])
    assert np.max(np.abs(occ - g["occ_after"]) / np.abs(g["occ_after"])) > 1e-3
