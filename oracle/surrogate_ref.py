"""FP64 numpy restatement of the spectral neural operator of
include/pmc_surrogate.h: forward pass, mean-squared-error loss and the full
gradient, for checking the device kernels.  Test infrastructure only.

The parameter layout, the Fourier basis D (cos/sin of |kx| < modes_x,
0 <= kz < modes_z on the nx x nz grid, G = D / S) and the tanh form of GELU
follow the header's definition."""
import numpy as np


def gelu(x):
    u = np.sqrt(2.0 / np.pi) * (x + 0.044715 * x ** 3)
    return 0.5 * x * (1.0 + np.tanh(u))


def gelu_grad(x):
    c = np.sqrt(2.0 / np.pi)
    u = c * (x + 0.044715 * x ** 3)
    th = np.tanh(u)
    return 0.5 * (1.0 + th) + 0.5 * x * (1.0 - th * th) * c * (1.0 + 3.0 * 0.044715 * x * x)


def basis(nx, nz, mx, mz):
    cols = []
    ix, iz = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    ix, iz = ix.reshape(-1), iz.reshape(-1)
    for kx in range(-(mx - 1), mx):
        for kz in range(mz):
            th = 2.0 * np.pi * (kx * ix / nx + kz * iz / nz)
            cols += [np.cos(th), np.sin(th)]
    return np.stack(cols, axis=1)


def unpack(p, cin, w, layers, m, h):
    o = 0

    def take(*shape):
        nonlocal o
        n = int(np.prod(shape))
        v = p[o:o + n].reshape(shape)
        o += n
        return v

    P, bp = take(cin, w), take(w)
    R, Wl, bl = [], [], []
    for _ in range(layers):
        R.append(take(m, 2 * w, 2 * w))
        Wl.append(take(w, w))
        bl.append(take(w))
    Q1, b1, Q2, b2 = take(w, h), take(h), take(h, 1), take(1)
    assert o == p.size
    return P, bp, R, Wl, bl, Q1, b1, Q2, b2


def loss_and_grad(p, x, t, nx, nz, cin, w, layers, mx, mz, h):
    """p: flat parameters; x: [B, S, cin]; t: [B, S]. Returns (loss, grad)."""
    p = p.astype(np.float64)
    x = x.astype(np.float64)
    t = t.astype(np.float64)
    S, M = nx * nz, (2 * mx - 1) * mz
    P, bp, R, Wl, bl, Q1, b1, Q2, b2 = unpack(p, cin, w, layers, M, h)
    D = basis(nx, nz, mx, mz)
    G = D / S
    B = x.shape[0]
    hs, pres, Ys = [x @ P + bp], [], []
    for l in range(layers):
        Y = np.einsum("sm,bsw->bmw", D, hs[l])                       # [B, 2M, W]
        v = Y.reshape(B, M, 2 * w)
        Z = np.einsum("bmi,mij->bmj", v, R[l]).reshape(B, 2 * M, w)
        pre = np.einsum("sm,bmw->bsw", G, Z) + hs[l] @ Wl[l] + bl[l]
        Ys.append(v)
        pres.append(pre)
        hs.append(gelu(pre) if l < layers - 1 else pre)
    preq = hs[-1] @ Q1 + b1
    aq = gelu(preq)
    y = (aq @ Q2)[..., 0] + b2[0]
    n = y.size
    loss = np.mean((y - t) ** 2)
    dy = 2.0 * (y - t) / n
    gQ2 = np.einsum("bsh,bs->h", aq, dy)[:, None]
    gb2 = np.array([dy.sum()])
    daq = dy[..., None] * Q2[:, 0]
    dpq = daq * gelu_grad(preq)
    gQ1 = np.einsum("bsw,bsh->wh", hs[-1], dpq)
    gb1 = dpq.sum(axis=(0, 1))
    dh = dpq @ Q1.T
    gR, gW, gb = [None] * layers, [None] * layers, [None] * layers
    for l in range(layers - 1, -1, -1):
        dpre = dh * (gelu_grad(pres[l]) if l < layers - 1 else 1.0)
        gb[l] = dpre.sum(axis=(0, 1))
        gW[l] = np.einsum("bsi,bsj->ij", hs[l], dpre)
        dskip = dpre @ Wl[l].T
        dZ = np.einsum("sm,bsw->bmw", G, dpre).reshape(B, M, 2 * w)
        gR[l] = np.einsum("bmi,bmj->mij", Ys[l], dZ)
        dY = np.einsum("bmj,mij->bmi", dZ, R[l]).reshape(B, 2 * M, w)
        dh = dskip + np.einsum("sm,bmw->bsw", D, dY)
    gP = np.einsum("bsc,bsw->cw", x, dh)
    gbp = dh.sum(axis=(0, 1))
    parts = [gP, gbp]
    for l in range(layers):
        parts += [gR[l], gW[l], gb[l]]
    parts += [gQ1, gb1, gQ2, gb2]
    return loss, np.concatenate([q.reshape(-1) for q in parts])


def forward(p, x, nx, nz, cin, w, layers, mx, mz, h):
    """Predictions [B, S] (FP64)."""
    S, M = nx * nz, (2 * mx - 1) * mz
    P, bp, R, Wl, bl, Q1, b1, Q2, b2 = unpack(p.astype(np.float64), cin, w, layers, M, h)
    D = basis(nx, nz, mx, mz)
    G = D / S
    x = x.astype(np.float64)
    B = x.shape[0]
    hcur = x @ P + bp
    for l in range(layers):
        v = np.einsum("sm,bsw->bmw", D, hcur).reshape(B, M, 2 * w)
        Z = np.einsum("bmi,mij->bmj", v, R[l]).reshape(B, 2 * M, w)
        pre = np.einsum("sm,bmw->bsw", G, Z) + hcur @ Wl[l] + bl[l]
        hcur = gelu(pre) if l < layers - 1 else pre
    return (gelu(hcur @ Q1 + b1) @ Q2)[..., 0] + b2[0]
