"""Float64 CPU reference of the three-dimensional spectral operator of
include/pmc_fno.h, written with torch.fft and torch autograd.  Test
infrastructure: the GPU operator's loss and gradient are checked against it.

It restates the architecture, not the device's separable cos/sin
contractions: the spectral convolution is torch.fft.rfftn with norm="ortho",
the kept modes (kx and ky in [0, m) U [n - m, n), kz in [0, m)) multiplied per
mode by a complex W x W matrix, and torch.fft.irfftn back (norm="ortho").
Parameters follow the flat layout of pmc_fno_create."""
import numpy as np
import torch


def unpack(params, shape):
    cin, w, h, layers, m = shape["in_channels"], shape["width"], shape["hidden"], shape["layers"], shape["modes"]
    M = 4 * m * m * m
    p = torch.as_tensor(np.asarray(params, dtype=np.float64))
    o = 0

    def take(n, *dims):
        nonlocal o
        t = p[o:o + n].reshape(*dims)
        o += n
        return t

    out = {"P": take(cin * w, cin, w), "bp": take(w, w), "R": [], "W": [], "b": []}
    for _ in range(layers):
        out["R"].append(take(M * 2 * w * w, M, 2, w, w))
        out["W"].append(take(w * w, w, w))
        out["b"].append(take(w, w))
    out["Q1"] = take(w * h, w, h)
    out["b1"] = take(h, h)
    out["Q2"] = take(h, h)
    out["b2"] = take(1, 1)
    assert o == p.numel()
    return out


def spectral(h, R, m):
    """K(h) for h [B, X, Y, Z, W] and complex weights R [M, 2, W, W]."""
    B, X, Y, Z, W = h.shape
    F = torch.fft.rfftn(h, dim=(1, 2, 3), norm="ortho")
    kx = list(range(m)) + list(range(X - m, X))
    ky = list(range(m)) + list(range(Y - m, Y))
    sel = F[:, kx][:, :, ky][:, :, :, :m]                      # [B, 2m, 2m, m, W]
    Wc = torch.complex(R[:, 0], R[:, 1]).reshape(2 * m, 2 * m, m, W, W)
    mixed = torch.einsum("bxyzc,xyzcd->bxyzd", sel, Wc)
    U = torch.zeros((B, X, Y, Z // 2 + 1, W), dtype=F.dtype)
    ix = torch.tensor(kx)[:, None]
    iy = torch.tensor(ky)[None, :]
    U[:, ix, iy, :m] = mixed
    return torch.fft.irfftn(U, s=(X, Y, Z), dim=(1, 2, 3), norm="ortho")


def gelu(x):
    return torch.nn.functional.gelu(x, approximate="tanh")


def forward(params, shape, x):
    """Prediction [B, S] for inputs x [B, S, Cin] (numpy or torch)."""
    q = unpack(params, shape) if not isinstance(params, dict) else params
    X, Y, Z = shape["nx"], shape["ny"], shape["nz"]
    x = torch.as_tensor(np.asarray(x, dtype=np.float64))
    B = x.shape[0]
    h = (x @ q["P"] + q["bp"]).reshape(B, X, Y, Z, -1)
    L = len(q["R"])
    for l in range(L):
        pre = spectral(h, q["R"][l], shape["modes"]) + h @ q["W"][l] + q["b"][l]
        h = gelu(pre) if l < L - 1 else pre
    y = gelu(h @ q["Q1"] + q["b1"]) @ q["Q2"] + q["b2"]
    return y.reshape(B, -1)


def loss_grad(params, shape, x, t, w):
    """Weighted mean-squared error sum w (y - t)^2 / sum w and its gradient (flat, float64)."""
    p = torch.tensor(np.asarray(params, dtype=np.float64), requires_grad=True)
    y = forward(unpack_torch(p, shape), shape, x)
    t = torch.as_tensor(np.asarray(t, dtype=np.float64))
    w = torch.as_tensor(np.asarray(w, dtype=np.float64))
    loss = (w * (y - t) ** 2).sum() / w.sum()
    loss.backward()
    return float(loss.detach()), p.grad.numpy().copy()


def unpack_torch(p, shape):
    """unpack() on a torch tensor that carries gradients."""
    cin, w, h, layers, m = shape["in_channels"], shape["width"], shape["hidden"], shape["layers"], shape["modes"]
    M = 4 * m * m * m
    o = 0

    def take(n, *dims):
        nonlocal o
        t = p[o:o + n].reshape(*dims)
        o += n
        return t

    out = {"P": take(cin * w, cin, w), "bp": take(w, w), "R": [], "W": [], "b": []}
    for _ in range(layers):
        out["R"].append(take(M * 2 * w * w, M, 2, w, w))
        out["W"].append(take(w * w, w, w))
        out["b"].append(take(w, w))
    out["Q1"] = take(w * h, w, h)
    out["b1"] = take(h, h)
    out["Q2"] = take(h, h)
    out["b2"] = take(1, 1)
    return out
