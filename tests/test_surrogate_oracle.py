"""The FP64 numpy restatement of the spectral neural operator (oracle/surrogate_ref.py)
against central finite differences of its own loss, and its Fourier basis
against the closed form, so the oracle that checks the device gradients is
itself pinned by calculus rather than by restating its formulas."""
import numpy as np

from oracle import surrogate_ref as ref

SHAPE = dict(nx=6, nz=4, cin=3, w=4, layers=2, mx=3, mz=2, h=5)


def sizes():
    s = SHAPE
    m = (2 * s["mx"] - 1) * s["mz"]
    n = s["cin"] * s["w"] + s["w"] + s["layers"] * (m * 4 * s["w"] ** 2 + s["w"] ** 2 + s["w"]) + s["w"] * s["h"] + 2 * s["h"] + 1
    return m, n


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    m, n = sizes()
    p = rng.normal(scale=0.3, size=n)
    S = SHAPE["nx"] * SHAPE["nz"]
    x = rng.normal(size=(2, S, SHAPE["cin"]))
    t = rng.normal(size=(2, S))
    loss, g = ref.loss_and_grad(p, x, t, **SHAPE)
    assert g.shape == (n,)
    idx = rng.choice(n, 60, replace=False)
    eps = 1e-6
    for i in idx:
        e = np.zeros(n)
        e[i] = eps
        lp, _ = ref.loss_and_grad(p + e, x, t, **SHAPE)
        lm, _ = ref.loss_and_grad(p - e, x, t, **SHAPE)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    np.testing.assert_allclose(ref.forward(p, x, **SHAPE), ref.forward(p, x, **SHAPE))
    assert loss == np.mean((ref.forward(p, x, **SHAPE) - t) ** 2)


def test_basis_is_the_real_fourier_basis():
    d = ref.basis(6, 4, 3, 2)
    assert d.shape == (24, 2 * 5 * 2)
    # kx = kz = 0: cos column all ones, sin column all zeros
    m0 = (3 - 1) * 2  # index of (kx=0, kz=0)
    np.testing.assert_allclose(d[:, 2 * m0], 1.0)
    np.testing.assert_allclose(d[:, 2 * m0 + 1], 0.0, atol=1e-15)
    # distinct (kx, kz >= 0) cos/sin columns are orthogonal on the grid
    gram = d.T @ d
    off = gram - np.diag(np.diag(gram))
    assert np.abs(off[2 * m0:, 2 * m0:]).max() < 1e-9
