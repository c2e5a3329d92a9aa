"""ctypes binding of the three-dimensional spectral neural operator
(include/pmc_fno.h): BF16 matrix-core forward, backward and Adam on the device.

Inputs are [batch][nx*ny*nz][in_channels] float32, channels last, point index
(x*ny + y)*nz + z; outputs [batch][nx*ny*nz]."""
import ctypes as C

import numpy as np

from . import runtime as rt


class Shape(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("in_channels", C.c_int32),
                ("width", C.c_int32), ("layers", C.c_int32), ("modes", C.c_int32), ("hidden", C.c_int32),
                ("max_batch", C.c_int32), ("seed", C.c_uint64)]


_bound = False


def _lib():
    global _bound
    L = rt.lib()
    if not _bound:
        P, f = C.c_void_p, C.POINTER(C.c_float)
        L.pmc_fno_create.argtypes = [C.POINTER(Shape), C.POINTER(P)]
        L.pmc_fno_param_count.restype = C.c_int64
        L.pmc_fno_param_count.argtypes = [P]
        L.pmc_fno_get_params.argtypes = [P, f]
        L.pmc_fno_set_params.argtypes = [P, f]
        L.pmc_fno_predict.argtypes = [P, C.c_int32, f, f]
        L.pmc_fno_loss_grad.argtypes = [P, C.c_int32, f, f, f, f, f]
        L.pmc_fno_train.argtypes = [P, C.c_int32, f, f, f, C.c_int32, C.c_int32, C.c_float, C.c_uint64, f]
        L.pmc_fno_destroy.argtypes = [P]
        L.pmc_fno_set_timing.argtypes = [P, C.c_int32]
        L.pmc_fno_gemm_stats.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                         C.c_int32]
        _bound = True
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Operator3d:
    """pmc_fno: spectral operator on an nx x ny x nz grid."""

    def __init__(self, nx, ny, nz, in_channels=3, width=32, layers=4, modes=12, hidden=128, max_batch=8,
                 seed=20261204):
        self.shape = Shape(nx, ny, nz, in_channels, width, layers, modes, hidden, max_batch, seed)
        h = C.c_void_p()
        rc = _lib().pmc_fno_create(C.byref(self.shape), C.byref(h))
        if rc:
            raise RuntimeError(f"pmc_fno_create failed: {rc}")
        self.h = h
        self.S = nx * ny * nz
        self.nparam = _lib().pmc_fno_param_count(h)

    def params(self):
        p = np.zeros(self.nparam, dtype=np.float32)
        if _lib().pmc_fno_get_params(self.h, _fp(p)):
            raise RuntimeError("pmc_fno_get_params failed")
        return p

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.nparam
        if _lib().pmc_fno_set_params(self.h, _fp(p)):
            raise RuntimeError("pmc_fno_set_params failed")

    def predict(self, x):
        x = _f32(x)
        out = np.zeros((x.shape[0], self.S), dtype=np.float32)
        rc = _lib().pmc_fno_predict(self.h, x.shape[0], _fp(x), _fp(out))
        if rc:
            raise RuntimeError(f"pmc_fno_predict failed: {rc}")
        return out

    def loss_grad(self, x, t, w):
        x, t, w = _f32(x), _f32(t), _f32(w)
        loss = C.c_float()
        g = np.zeros(self.nparam, dtype=np.float32)
        rc = _lib().pmc_fno_loss_grad(self.h, x.shape[0], _fp(x), _fp(t), _fp(w), C.byref(loss), _fp(g))
        if rc:
            raise RuntimeError(f"pmc_fno_loss_grad failed: {rc}")
        return loss.value, g

    def train(self, x, t, w, epochs, batch=4, lr=1e-3, seed=1):
        """x, t, w are kept referenced: the device copy of the data set is reused
        while the same arrays are passed again."""
        self._data = (_f32(x), _f32(t), _f32(w))
        x, t, w = self._data
        losses = np.zeros(epochs, dtype=np.float32)
        rc = _lib().pmc_fno_train(self.h, x.shape[0], _fp(x), _fp(t), _fp(w), epochs, batch, lr, seed, _fp(losses))
        if rc:
            raise RuntimeError(f"pmc_fno_train failed: {rc}")
        return losses

    def set_timing(self, on=True):
        if _lib().pmc_fno_set_timing(self.h, 1 if on else 0):
            raise RuntimeError("pmc_fno_set_timing failed")

    def gemm_stats(self, reset=False):
        """(flops, launches, ms) of the operator's GEMM launches since the last reset."""
        f, n, ms = C.c_double(), C.c_int64(), C.c_double()
        if _lib().pmc_fno_gemm_stats(self.h, C.byref(f), C.byref(n), C.byref(ms), 1 if reset else 0):
            raise RuntimeError("pmc_fno_gemm_stats failed")
        return f.value, n.value, ms.value

    def close(self):
        if getattr(self, "h", None):
            _lib().pmc_fno_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
