"""Device kernels of the spectral neural operator: the BF16 matrix-core GEMM
against numpy on BF16-rounded operands (every transpose combination, ragged
sizes, batches, split k), the loss and full gradient against the FP64 oracle
(oracle/surrogate_ref.py), and a short training run."""
import numpy as np
import pytest

from oracle import surrogate_ref as ref
from phonomc_amd import surrogate as sg

pytestmark = pytest.mark.gpu


def bf16(a):
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("m,n,k,batch", [(1, 1, 1, 1), (17, 33, 65, 1), (64, 64, 16, 3), (130, 70, 231, 2),
                                         (32, 48, 4003, 2), (1, 64, 1000, 1)])  # the last two split k
def test_gemm_matches_numpy(m, n, k, batch, ta, tb):
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    a = rng.normal(size=(batch, k, m) if ta else (batch, m, k)).astype(np.float32)
    b = rng.normal(size=(batch, n, k) if tb else (batch, k, n)).astype(np.float32)
    c0 = rng.normal(size=(batch, m, n)).astype(np.float32)
    out = sg.gemm(a, b, c0, ta=bool(ta), tb=bool(tb), alpha=0.75, beta=0.5)
    A = bf16(a).astype(np.float64)
    B = bf16(b).astype(np.float64)
    A = np.swapaxes(A, 1, 2) if ta else A
    B = np.swapaxes(B, 1, 2) if tb else B
    want = 0.75 * (A @ B) + 0.5 * c0
    np.testing.assert_allclose(out, want, rtol=0, atol=1e-5 * max(1, k) ** 0.5 * 4)


SHAPE = dict(nx=10, nz=4, cin=5, w=8, layers=2, mx=3, mz=2, h=16)


def make_op():
    s = SHAPE
    return sg.Operator(s["nx"], s["nz"], in_channels=s["cin"], width=s["w"], layers=s["layers"], modes_x=s["mx"],
                       modes_z=s["mz"], hidden=s["h"], max_batch=8, seed=11)


def test_basis_matches_oracle():
    op = make_op()
    np.testing.assert_allclose(op.basis(), ref.basis(10, 4, 3, 2), atol=1e-6)


def test_loss_and_gradient_match_fp64_oracle():
    """BF16 products with FP32 sums: the gradient agrees with the FP64 oracle in
    direction (cosine > 0.999) and to 3 % in norm; the loss to 1 %."""
    op = make_op()
    rng = np.random.default_rng(3)
    S = 40
    x = rng.normal(size=(4, S, 5)).astype(np.float32)
    t = rng.normal(size=(4, S)).astype(np.float32)
    p = op.params()
    loss, g = op.loss_grad(x, t)
    lref, gref = ref.loss_and_grad(p, x, t, **SHAPE)
    assert loss == pytest.approx(lref, rel=1e-2)
    cos = float(g @ gref / (np.linalg.norm(g) * np.linalg.norm(gref)))
    rel = float(np.linalg.norm(g - gref) / np.linalg.norm(gref))
    print(f"loss {loss:.6f} vs {lref:.6f}; grad cosine {cos:.6f}, relative L2 {rel:.3e}")
    assert cos > 0.999 and rel < 3e-2
    np.testing.assert_allclose(op.predict(x), ref.forward(p, x, **SHAPE), rtol=0, atol=5e-2 * np.abs(t).max())


def test_training_reduces_the_loss():
    op = make_op()
    rng = np.random.default_rng(5)
    S = 40
    x = rng.normal(size=(64, S, 5)).astype(np.float32)
    t = (np.tanh(x[..., 0]) + 0.5 * x[..., 1]).astype(np.float32)
    losses = op.train(x, t, epochs=40, batch=8, lr=3e-3)
    print("losses", losses[0], losses[-1])
    assert losses[-1] < 0.3 * losses[0]
