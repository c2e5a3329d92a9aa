"""The device 3-D spectral operator (include/pmc_fno.h) against the float64
torch reference (oracle/fno_ref.py): the same parameters, inputs, targets and
weights give the same prediction, loss and gradient up to the BF16 rounding of
the matrix-core operands (the bar written in each assertion), and device
training lowers the loss."""
import numpy as np
import pytest

from phonomc_amd import fno

pytestmark = pytest.mark.gpu
SHAPE = dict(nx=16, ny=8, nz=8, in_channels=3, width=8, layers=2, modes=2, hidden=16)


def make(max_batch=2, seed=5):
    s = SHAPE
    return fno.Operator3d(s["nx"], s["ny"], s["nz"], s["in_channels"], s["width"], s["layers"], s["modes"],
                          s["hidden"], max_batch=max_batch, seed=seed)


def data(rng, b):
    S = SHAPE["nx"] * SHAPE["ny"] * SHAPE["nz"]
    x = rng.normal(size=(b, S, 3)).astype(np.float32)
    t = rng.normal(size=(b, S)).astype(np.float32)
    w = (rng.random((b, S)) < 0.6).astype(np.float32)
    return x, t, w


def test_prediction_loss_and_gradient_match_reference():
    from oracle import fno_ref

    rng = np.random.default_rng(11)
    op = make()
    p = op.params()
    p[:] = rng.normal(scale=0.2, size=p.size).astype(np.float32)  # well away from initialization
    op.set_params(p)
    x, t, w = data(rng, 2)
    y = op.predict(x)
    y_ref = fno_ref.forward(p.astype(np.float64), SHAPE, x.astype(np.float64)).numpy()
    rel = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    loss, g = op.loss_grad(x, t, w)
    loss_ref, g_ref = fno_ref.loss_grad(p.astype(np.float64), SHAPE, x, t, w)
    cos = float(g @ g_ref / np.linalg.norm(g) / np.linalg.norm(g_ref))
    grel = float(np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref))
    print(f"prediction rel. L2 {rel:.2e}; loss {loss:.6f} vs {loss_ref:.6f}; gradient cosine {cos:.6f}, "
          f"rel. L2 {grel:.2e}")
    assert rel < 2e-2                      # BF16 operands (8-bit mantissa) through 2 layers of 9 products
    assert loss == pytest.approx(loss_ref, rel=2e-2)
    assert cos > 0.999 and grel < 5e-2


def test_device_training_lowers_the_loss():
    rng = np.random.default_rng(12)
    S = SHAPE["nx"] * SHAPE["ny"] * SHAPE["nz"]
    x, _, _ = data(rng, 8)
    # A smooth target the operator can represent: a low-mode function of the input field.
    t = np.tanh(x[:, :, 0] + 0.5 * x[:, :, 1]).astype(np.float32)
    w = np.ones((8, S), dtype=np.float32)
    op = make(max_batch=4)
    before, _ = op.loss_grad(x[:4], t[:4], w[:4])
    losses = op.train(x, t, w, epochs=30, batch=4, lr=3e-3)
    after, _ = op.loss_grad(x[:4], t[:4], w[:4])
    print(f"loss {before:.4f} -> {after:.4f}; epochs {losses[0]:.4f} .. {losses[-1]:.4f}")
    assert after < 0.5 * before and losses[-1] < losses[0]
