"""ctypes binding of the spectral neural operator (include/pmc_surrogate.h) and
the feature layout of the film warm start.

Inputs per cell (channels): heat-source power density (normalized), solid mask,
reservoir/substrate temperature (normalized), and the cell's x and z position in
[0, 1).  The target is the steady cell temperature minus the reservoir
temperature, in kelvin.
"""
import ctypes as C

import numpy as np

from . import runtime as rt


class Shape(C.Structure):
    _fields_ = [("nx", C.c_int32), ("nz", C.c_int32), ("in_channels", C.c_int32), ("width", C.c_int32),
                ("layers", C.c_int32), ("modes_x", C.c_int32), ("modes_z", C.c_int32), ("hidden", C.c_int32),
                ("max_batch", C.c_int32), ("seed", C.c_uint64)]


_bound = False


def _lib():
    global _bound
    L = rt.lib()
    if not _bound:
        P, f = C.c_void_p, C.POINTER(C.c_float)
        L.pmc_nn_create.argtypes = [C.POINTER(Shape), C.POINTER(P)]
        L.pmc_nn_param_count.restype = C.c_int64
        L.pmc_nn_param_count.argtypes = [P]
        L.pmc_nn_get_params.argtypes = [P, f]
        L.pmc_nn_set_params.argtypes = [P, f]
        L.pmc_nn_basis.argtypes = [P, f]
        L.pmc_nn_predict.argtypes = [P, C.c_int32, f, f]
        L.pmc_nn_loss_grad.argtypes = [P, C.c_int32, f, f, f, f]
        L.pmc_nn_train.argtypes = [P, C.c_int32, f, f, C.c_int32, C.c_int32, C.c_float, C.c_uint64, f]
        L.pmc_nn_destroy.argtypes = [P]
        L.pmc_nn_set_timing.argtypes = [P, C.c_int32]
        L.pmc_nn_gemm_stats.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                        C.c_int32]
        L.pmc_nn_gemm.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_float, f, C.c_int32, C.c_int64,
                                  C.c_int32, f, C.c_int32, C.c_int64, C.c_int32, C.c_float, f, C.c_int32, C.c_int64]
        _bound = True
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def gemm(a, b, c=None, ta=False, tb=False, alpha=1.0, beta=0.0):
    """Batched C = alpha op(A) op(B) + beta C on the device (BF16 products, FP32 sums).
    a: [batch, rows, cols] (or 2-D), same for b; returns [batch, m, n]."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    a3, b3 = (a[None] if a.ndim == 2 else a), (b[None] if b.ndim == 2 else b)
    batch = max(a3.shape[0], b3.shape[0])
    m, k = (a3.shape[2], a3.shape[1]) if ta else (a3.shape[1], a3.shape[2])
    n = b3.shape[1] if tb else b3.shape[2]
    out = np.zeros((batch, m, n), dtype=np.float32) if c is None else np.ascontiguousarray(c, dtype=np.float32).copy()
    sa = 0 if a3.shape[0] == 1 else a3.shape[1] * a3.shape[2]
    sb = 0 if b3.shape[0] == 1 else b3.shape[1] * b3.shape[2]
    rc = _lib().pmc_nn_gemm(m, n, k, batch, alpha, _fp(a3), a3.shape[2], sa, int(ta), _fp(b3), b3.shape[2], sb,
                            int(tb), beta, _fp(out), n, m * n)
    if rc:
        raise RuntimeError(f"pmc_nn_gemm failed: {rc}")
    return out


class Operator:
    """Spectral neural operator on an nx x nz cell grid (pmc_nn)."""

    def __init__(self, nx, nz, in_channels=5, width=32, layers=4, modes_x=12, modes_z=5, hidden=128, max_batch=64,
                 seed=20261203):
        self.shape = Shape(nx, nz, in_channels, width, layers, modes_x, modes_z, hidden, max_batch, seed)
        h = C.c_void_p()
        rc = _lib().pmc_nn_create(C.byref(self.shape), C.byref(h))
        if rc:
            raise RuntimeError(f"pmc_nn_create failed: {rc}")
        self.h = h
        self.S = nx * nz
        self.nparam = _lib().pmc_nn_param_count(h)

    def params(self):
        p = np.zeros(self.nparam, dtype=np.float32)
        _lib().pmc_nn_get_params(self.h, _fp(p))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        _lib().pmc_nn_set_params(self.h, _fp(p))

    def basis(self):
        m = (2 * self.shape.modes_x - 1) * self.shape.modes_z
        d = np.zeros((self.S, 2 * m), dtype=np.float32)
        _lib().pmc_nn_basis(self.h, _fp(d))
        return d

    def predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros((x.shape[0], self.S), dtype=np.float32)
        rc = _lib().pmc_nn_predict(self.h, x.shape[0], _fp(x), _fp(out))
        if rc:
            raise RuntimeError(f"pmc_nn_predict failed: {rc}")
        return out

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float()
        g = np.zeros(self.nparam, dtype=np.float32)
        rc = _lib().pmc_nn_loss_grad(self.h, x.shape[0], _fp(x), _fp(t), C.byref(loss), _fp(g))
        if rc:
            raise RuntimeError(f"pmc_nn_loss_grad failed: {rc}")
        return loss.value, g

    def train(self, x, t, epochs, batch=32, lr=1e-3, seed=1):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        losses = np.zeros(epochs, dtype=np.float32)
        rc = _lib().pmc_nn_train(self.h, x.shape[0], _fp(x), _fp(t), epochs, batch, lr, seed, _fp(losses))
        if rc:
            raise RuntimeError(f"pmc_nn_train failed: {rc}")
        return losses

    def set_timing(self, on=True):
        if _lib().pmc_nn_set_timing(self.h, 1 if on else 0):
            raise RuntimeError("pmc_nn_set_timing failed")

    def gemm_stats(self, reset=False):
        """(flops, launches, ms) of the operator's GEMM launches since the last reset."""
        f, n, ms = C.c_double(), C.c_int64(), C.c_double()
        if _lib().pmc_nn_gemm_stats(self.h, C.byref(f), C.byref(n), C.byref(ms), 1 if reset else 0):
            raise RuntimeError("pmc_nn_gemm_stats failed")
        return f.value, n.value, ms.value

    def close(self):
        if getattr(self, "h", None):
            _lib().pmc_nn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
