"""Host side of the FinFET warm start (phonomc_amd/warmstart3d.py): the
mirror quadrant of the 64 x 32 x 32 FinFET grid, folding and unfolding, the
16-bit data-set format, the power cap and the per-case loss weights."""
import os
import tempfile

import numpy as np
import pytest

from phonomc_amd import warmstart3d as w3


@pytest.fixture(scope="module")
def geom():
    return w3.Geometry(w3.problem_of(w3.sample_cases(1, 3)[0], 20000, 1))


def test_quadrant_covers_the_solid_with_its_mirror_images(geom):
    quad = w3.quadrant(geom)
    assert quad.shape[0] == 4
    assert 4 * quad.shape[1] == geom.n_cells        # no voxel on a mirror plane (even grid)
    assert len(np.unique(quad)) == geom.n_cells
    assert np.all(geom.mask[quad.ravel()] == 1.0)


def test_fold_unfold(geom):
    quad = w3.quadrant(geom)
    rng = np.random.default_rng(1)
    q = rng.normal(size=(3, quad.shape[1]))
    full = w3.unfold(q, quad)
    np.testing.assert_allclose(w3.fold(full, quad), q, rtol=1e-6)    # identity on mirror-symmetric fields
    g = rng.normal(size=(2, int(np.prod(w3.GRID)))) * geom.mask
    f = w3.fold(g, quad)                                               # mean of the four images
    np.testing.assert_allclose(f, g[:, quad].mean(axis=1))
    X, Y, Z = w3.GRID
    s = w3.unfold(f, quad).reshape(2, X, Y, Z)
    np.testing.assert_allclose(s, s[:, ::-1, :, :])
    np.testing.assert_allclose(s, s[:, :, ::-1, :])


def test_source_feature_is_mirror_symmetric(geom):
    p = dict(height=37.0, sigma=6.0, power=3e20, t_bg=310.0)
    f = w3.features(w3.problem_of(p, 20000, 1), p, geom)
    X, Y, Z = w3.GRID
    d = f[:, 0].reshape(X, Y, Z)
    np.testing.assert_array_equal(d, d[::-1])
    np.testing.assert_array_equal(d, d[:, ::-1])
    assert f[:, 0].max() == 1.0 and np.all(f[geom.mask == 0] == 0)
    np.testing.assert_allclose(f[geom.mask > 0, 2], (310.0 - w3.T_MID) / w3.T_HALF)


def test_run_power_is_capped_in_total_power():
    small = dict(height=50.0, sigma=3.0, power=1e19, t_bg=300.0)   # about 5 uW: runs as sampled
    big = dict(height=50.0, sigma=15.0, power=1e21, t_bg=300.0)    # about 53 mW: capped
    assert w3.run_power_density(small) == small["power"]
    total = w3.run_power_density(big) * (2 * np.pi) ** 1.5 * (15e-9) ** 3
    assert total == pytest.approx(w3.P_CAP_W, rel=1e-12)


def test_dataset_round_trip(geom):
    """pack_dataset stores each case's rise per microwatt, mirror-averaged, at 16
    bits: loading gives the folded field back within half a quantization step."""
    rng = np.random.default_rng(2)
    cases = w3.sample_cases(5, 11)
    quad = w3.quadrant(geom)
    sym = w3.unfold(rng.gamma(2.0, size=(5, quad.shape[1])), quad)   # symmetric rise per microwatt (grid)
    watts = rng.uniform(1e-5, 4e-4, 5)
    rise = np.stack([w3.Geometry.to_cells(geom, sym[k]) for k in range(5)]) * (watts[:, None] / w3.Y_UNIT_W)
    tmp = tempfile.mkdtemp()
    raw = os.path.join(tmp, "raw.npz")
    np.savez(raw, params=np.array([[c[k] for k in w3.PARAM_KEYS] for c in cases]), watts=watts,
             rise=rise.astype(np.float32))
    out = os.path.join(tmp, "packed.npz")
    problem = w3.problem_of(cases[0], 20000, 1)
    w3.pack_dataset([raw], out, problem)
    got_cases, per_uw, g2 = w3.load_dataset(out, problem)
    assert got_cases == [{k: pytest.approx(c[k]) for k in w3.PARAM_KEYS} for c in cases]
    step = (sym.max(axis=1) - sym.min(axis=1)) / 65535.0
    err = np.abs(per_uw - sym * geom.mask).max(axis=1)
    assert np.all(err <= 0.5 * step + 1e-6 * sym.max(axis=1))


def test_relative_weights_equalize_cases():
    mask = np.zeros((2, 10), dtype=np.float32)
    mask[:, :6] = 1
    y = np.zeros((2, 10), dtype=np.float32)
    y[0, :6] = 1.0
    y[1, :6] = 100.0
    w = w3.relative_weights(y, mask)
    assert np.all(w[:, 6:] == 0)
    np.testing.assert_allclose((w * y ** 2).sum(axis=1), [6.0, 6.0], rtol=1e-6)
