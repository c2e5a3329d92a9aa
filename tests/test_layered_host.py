"""Layered materials (a Ge-like layer on a Si-like one along z) on the host and
in the oracle, against the reference's own outputs (tests/golden/interface_c1l.npz,
material_ge16.npz, written by make_golden.py --layered from the reference sources):

* per-material carrier volumes equal the ones the reference's run summary
  prints (17 digits) and the energy weights equal BackgroundCache's
  (include/solver/BackgroundCache.h:13-19) from the reference's own active
  counts and normalizations;
* the cell materials follow LayeredMaterialLayout::material_at of the cell
  centres (src/solver/LayeredMaterialLayout.cpp:23-31);
* the Ge-like table's energy/temperature lookups are bit-exact with the
  reference's PhononMaterial, and the layered view concatenates the single
  material tables unchanged;
* the oracle's DMM event reproduces InterfaceInteraction::interact
  (src/solver/InterfaceInteraction.cpp:47-154) on the uniforms the reference
  consumed: action, outgoing material and mode, occupation and velocity;
* a planted mistake in the oracle's remap (dropping the frequency ratio) is
  caught by the same vectors.
"""
import math

import numpy as np
import pytest

from oracle import oracle_ctypes as O
from phonomc_amd import runtime as rt


@pytest.fixture(scope="module")
def layered():
    return rt.Problem(rt.config_text("c1l_layered.toml", **{"simulation.particle_count": 20000}),
                      input_directory=rt.CONFIGS, mode_sampling=1)


def test_layout_volumes_and_weights_match_reference(layered, golden):
    g = golden("interface_c1l.npz")
    lv = layered.layers()
    assert lv.count == 2 and lv.axis == 2
    vol = rt.arr(lv.spatial_volume, lv.count)
    np.testing.assert_allclose(vol, g["carrier_volumes"], rtol=1e-15)
    s, m, geo, i, rough, res = layered.views()
    np.testing.assert_allclose(m.energy_weight, g["energy_weights"], rtol=1e-14)
    mb = rt.arr(lv.mode_begin, lv.count + 1, np.int32)
    assert np.array_equal(np.diff(mb), g["active_counts"])
    assert m.n_active == mb[-1]
    centers, _ = layered.cells()
    planes = rt.arr(lv.planes, lv.count - 1)
    expect = np.searchsorted(planes, centers[:, 2], side="right")
    assert np.array_equal(rt.arr(lv.cell_material, len(centers), np.int32), expect)
    assert list(expect) == [0] * 10 + [1] * 10
    assert lv.band_width == 2.0 * math.pi * float(g["bin_thz"])


def test_ge_table_lookups_bit_exact(golden):
    g = golden("material_ge16.npz")
    p = rt.Problem(rt.config_text("c1_cross_plane.toml", **{"io.material_folder": '"synthetic:ge16"'}),
                   mode_sampling=1)
    L = rt.lib()
    s, m, geo, i, rough, res = p.views()
    assert np.array_equal(rt.arr(m.active_flat, m.n_active, np.int32), g["active_flat"])
    assert p.scalars()["energy_normalization"] == float(g["normalization"])
    for e, t in zip(g["lut_e"][::7], g["lut_t"][::7]):
        assert L.pmcs_problem_temperature_from_energy(p.h, e) == t
    for t, e in zip(g["e_t"], g["e_of_t"]):
        assert L.pmcs_problem_energy_from_temperature(p.h, t) == e


def test_layered_view_concatenates_single_material_tables(layered):
    s, m, geo, i, rough, res = layered.views()
    lv = layered.layers()
    lb = rt.arr(lv.lut_begin, lv.count + 1, np.int32)
    lut_t, lut_e = rt.arr(m.lut_t, m.lut_n), rt.arr(m.lut_e, m.lut_n)
    ebg = rt.arr(lv.background_energy_density, lv.count)
    mb = rt.arr(lv.mode_begin, lv.count + 1, np.int32)
    omega, neq = rt.arr(m.omega, m.n_active), rt.arr(m.neq_background, m.n_active)
    for k, folder in enumerate(['"synthetic:si16"', '"synthetic:ge16"']):
        single = rt.Problem(rt.config_text("c1_cross_plane.toml", **{"io.material_folder": folder}), mode_sampling=1)
        s1, m1, *_ = single.views()
        assert np.array_equal(lut_t[lb[k]:lb[k + 1]], rt.arr(m1.lut_t, m1.lut_n))
        assert np.array_equal(lut_e[lb[k]:lb[k + 1]], rt.arr(m1.lut_e, m1.lut_n))
        assert ebg[k] == m1.background_energy_density
        assert np.array_equal(omega[mb[k]:mb[k + 1]], rt.arr(m1.omega, m1.n_active))
        assert np.array_equal(neq[mb[k]:mb[k + 1]], rt.arr(m1.neq_background, m1.n_active))


def dmm_mismatches(sim, g, mb):
    bad = []
    for k in range(len(g["material"])):
        act, mode, occ, vel = sim.interface_dmm(mb[g["material"][k]] + g["active"][k], g["velocity"][k],
                                               g["occupation"][k], g["next_material"][k], g["uniforms"][k])
        ra = int(g["out_action"][k])
        src = mb[g["material"][k]] + g["active"][k]
        # A carrier whose velocity is mirrored against its mode can reflect into
        # that same mode: the state equals an axis mirror's, which is what the
        # reference's outcome records (same mode, axis velocity flipped).
        ok = act == ra or (ra == -1 and act == 0 and mode == src)
        ok = ok and mode == mb[g["out_material"][k]] + g["out_active"][k]
        ok = ok and occ == pytest.approx(g["out_occupation"][k], rel=1e-13, abs=0)
        ok = ok and np.array_equal(vel, g["out_velocity"][k])
        if not ok:
            bad.append(k)
    return bad


def test_interface_dmm_matches_reference(layered, golden):
    g = golden("interface_c1l.npz")
    lv = layered.layers()
    mb = rt.arr(lv.mode_begin, lv.count + 1, np.int32)
    sim = O.Sim(layered)
    assert dmm_mismatches(sim, g, mb) == []
    actions = set(np.unique(g["out_action"]))
    assert actions == {-1, 0, 1}  # transmissions, reflections into new modes and mirrors all covered
    # Mirrored carriers (velocity flipped against their mode) are among the inputs.
    s, m, *_ = layered.views()
    vel = rt.arr(m.velocity, 3 * m.n_active).reshape(-1, 3)
    own = vel[mb[g["material"]] + g["active"], 2]
    assert np.sum(np.sign(own) != np.sign(g["velocity"][:, 2])) > 100


def test_planted_remap_mistake_is_caught(layered, golden):
    """The same vectors reject an oracle whose occupation remap drops the
    frequency ratio (new deviation = old deviation)."""
    g = golden("interface_c1l.npz")
    lv = layered.layers()
    mb = rt.arr(lv.mode_begin, lv.count + 1, np.int32)
    s, m, *_ = layered.views()
    omega, neq = rt.arr(m.omega, m.n_active), rt.arr(m.neq_background, m.n_active)
    wrong = 0
    for k in np.where(g["out_action"] >= 0)[0][:400]:
        src = mb[g["material"][k]] + g["active"][k]
        dst = mb[g["out_material"][k]] + g["out_active"][k]
        planted = neq[dst] + (g["occupation"][k] - neq[src])  # ratio omega/omega' dropped
        right = neq[dst] + omega[src] * (g["occupation"][k] - neq[src]) / omega[dst]
        assert right == pytest.approx(g["out_occupation"][k], rel=1e-13)
        wrong += planted != pytest.approx(g["out_occupation"][k], rel=1e-13)
    assert wrong > 300


@pytest.mark.parametrize("over,msg", [
    ({"materials.interface_positions": "[49]"}, "materials.interface_positions must coincide with grid-cell boundaries."),
    ({"materials.interface_positions": "[100]"}, "materials.interface_positions must lie strictly inside the domain"),
    ({"materials.interface_positions": "[30, 60]"}, "must contain exactly folders.size()-1 values."),
    ({"materials.interface_frequency_bin_thz": "0"}, "materials.interface_frequency_bin_thz must be finite and positive."),
    ({"materials.interface_amm_fraction": "0.5"}, "DMM cannot have nonzero AMM fraction."),
    ({"materials.interface_model": '"amm"'}, "AMM/MMM requires one mass density per material layer."),
    ({"materials.interface_axis": '"x"', "simulation.grid_xyz": "[2, 1, 20]"},
     "Periodic boundaries normal to materials.interface_axis are not supported"),
])
def test_layered_config_rejections_use_reference_messages(over, msg):
    """The messages (and, where several apply, the first one) the reference's
    own solver setup gives for these inputs: ConfigValidation's material
    checks, then LayeredMaterialLayout::validated_cell_materials."""
    with pytest.raises(RuntimeError) as e:
        rt.Problem(rt.config_text("c1l_layered.toml", **over), input_directory=rt.CONFIGS)
    assert msg in str(e.value)


def test_oracle_layered_energy_balance(layered):
    """The oracle's layered film conserves energy per step to round-off and
    keeps every carrier's mode in the material of its position."""
    sim = O.Sim(layered)
    sim.initialize()
    lv = layered.layers()
    mb = rt.arr(lv.mode_begin, lv.count + 1, np.int32)
    e0 = sim.ledger()["thermal_energy_ev"]
    sim.step(5)
    led = sim.ledger()
    assert led["interface_events"] > led["interface_transmitted"] > 0
    c = sim.carriers()
    mat = np.searchsorted(mb[1:], c["mode"], side="right")
    side = (c["position"][:, 2] >= rt.arr(lv.planes, 1)[0]).astype(int)
    assert np.array_equal(mat, side)
    balance = e0 + led["total_injected_energy_ev"] - led["total_absorbed_energy_ev"] - led["thermal_energy_ev"]
    assert abs(balance) < 1e-9 * max(1.0, abs(e0))
