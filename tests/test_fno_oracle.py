"""The float64 reference of the 3-D spectral operator (oracle/fno_ref.py) pinned
by closed forms and finite differences, before it judges the device operator:
- with identity mode matrices the spectral convolution passes a kept Fourier
  mode unchanged and removes one outside the kept set (a dropped or misplaced
  mode, or a wrong normalization, fails);
- its autograd gradient equals central finite differences of its own loss."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
from oracle import fno_ref  # noqa: E402

SHAPE = dict(nx=12, ny=8, nz=8, in_channels=3, width=4, layers=2, modes=2, hidden=6)


def n_params(s):
    M = 4 * s["modes"] ** 3
    w, h = s["width"], s["hidden"]
    return s["in_channels"] * w + w + s["layers"] * (M * 2 * w * w + w * w + w) + w * h + h + h + 1


def identity_weights(m, w):
    R = torch.zeros((4 * m ** 3, 2, w, w), dtype=torch.float64)
    R[:, 0] = torch.eye(w, dtype=torch.float64)
    return R


@pytest.mark.parametrize("k,kept", [((1, -1, 1), True), ((0, 1, 0), True), ((-2, 0, 1), True),
                                    ((3, 0, 0), False), ((0, 0, 2), False), ((1, 4, 1), False)])
def test_spectral_convolution_passes_kept_modes_only(k, kept):
    X, Y, Z, W, m = 12, 8, 8, 3, 2
    x, y, z = np.meshgrid(np.arange(X), np.arange(Y), np.arange(Z), indexing="ij")
    phase = 2 * np.pi * (k[0] * x / X + k[1] * y / Y + k[2] * z / Z)
    amp = np.array([1.0, -0.5, 2.0])
    h = torch.tensor((np.cos(phase + 0.3)[..., None] * amp)[None])
    out = fno_ref.spectral(h, identity_weights(m, W), m).numpy()
    if kept:
        np.testing.assert_allclose(out, h.numpy(), atol=1e-12)
    else:
        assert np.abs(out).max() < 1e-12


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rng.normal(scale=0.3, size=n_params(SHAPE))
    S = SHAPE["nx"] * SHAPE["ny"] * SHAPE["nz"]
    x = rng.normal(size=(2, S, 3))
    t = rng.normal(size=(2, S))
    w = (rng.random((2, S)) < 0.6).astype(float)
    loss, g = fno_ref.loss_grad(p, SHAPE, x, t, w)
    for i in rng.choice(p.size, 12, replace=False):
        e = np.zeros_like(p)
        e[i] = 1e-6
        lp, _ = fno_ref.loss_grad(p + e, SHAPE, x, t, w)
        lm, _ = fno_ref.loss_grad(p - e, SHAPE, x, t, w)
        assert g[i] == pytest.approx((lp - lm) / 2e-6, rel=1e-5, abs=1e-9)
