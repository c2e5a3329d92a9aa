"""pmc_nn_gemm_noalloc: the surrogate GEMM on caller-owned device arrays, with
its split-k partial products in the library's fixed per-GPU buffer instead of
allocated memory.  Checked through the C ABI against pmc_nn_gemm (host arrays,
library-allocated workspace) on two calls in a row, the second smaller, both
with split k (so the fixed buffer is used and reused), and against numpy on
BF16-rounded operands."""
import ctypes as C

import numpy as np
import pytest

from phonomc_amd import runtime as rt

pytestmark = pytest.mark.gpu


def _hip():
    # HIP entry points as the library resolves them: a process that imports
    # torch after loading the library holds a second HIP runtime (torch's
    # bundled copy) whose null stream and device synchronization know nothing
    # of the library's launches, and a bare "libamdhip64.so" may name it.
    h = C.CDLL(rt.BUILD + "/libpmc_device.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def _lib():
    L = C.CDLL(rt.BUILD + "/libpmc_device.so")
    f = C.c_float
    P = C.POINTER(f)
    i32, i64 = C.c_int32, C.c_int64
    L.pmc_nn_gemm.argtypes = [i32, i32, i32, i32, f, P, i32, i64, i32, P, i32, i64, i32, f, P, i32, i64]
    L.pmc_nn_gemm_noalloc.argtypes = [i32, i32, i32, i32, f, C.c_void_p, i32, i64, i32, C.c_void_p, i32, i64, i32, f,
                                      C.c_void_p, i32, i64, C.c_void_p]
    return L


def _bf16(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def _case(rng, m, n, k, batch, ta, tb):
    a = rng.normal(size=(batch, k, m) if ta else (batch, m, k)).astype(np.float32)
    b = rng.normal(size=(batch, n, k) if tb else (batch, k, n)).astype(np.float32)
    c = rng.normal(size=(batch, m, n)).astype(np.float32)
    return a, b, c


def _reference(a, b, c, ta, tb, alpha, beta):
    A = _bf16(a).astype(np.float64)
    B = _bf16(b).astype(np.float64)
    A = np.transpose(A, (0, 2, 1)) if ta else A
    B = np.transpose(B, (0, 2, 1)) if tb else B
    return alpha * A @ B + beta * c


def _run_both(hip, L, a, b, c, ta, tb, alpha, beta):
    batch, m, n = c.shape
    k = a.shape[1] if ta else a.shape[2]
    lda, ldb = a.shape[2], b.shape[2]
    fp = lambda x: x.ctypes.data_as(C.POINTER(C.c_float))
    host = c.copy()
    assert L.pmc_nn_gemm(m, n, k, batch, alpha, fp(a), lda, a[0].size, ta, fp(b), ldb, b[0].size, tb, beta, fp(host),
                         n, m * n) == 0
    dev = []
    for x in (a, b, c):
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), x.nbytes) == 0
        assert hip.hipMemcpy(p, x.ctypes.data, x.nbytes, 1) == 0
        dev.append(p)
    rc = L.pmc_nn_gemm_noalloc(m, n, k, batch, alpha, dev[0], lda, a[0].size, ta, dev[1], ldb, b[0].size, tb, beta,
                               dev[2], n, m * n, None)
    assert rc == 0
    assert hip.hipDeviceSynchronize() == 0
    out = np.empty_like(c)
    assert hip.hipMemcpy(out.ctypes.data, dev[2], out.nbytes, 2) == 0
    for p in dev:
        hip.hipFree(p)
    return host, out


def test_noalloc_gemm_matches_allocating_gemm_twice():
    hip, L = _hip(), _lib()
    rng = np.random.default_rng(7)
    # first call: split k over 16 partial products; second, smaller: 5
    for (m, n, k, batch, ta, tb) in [(96, 80, 2048, 3, 0, 1), (40, 24, 640, 2, 1, 0)]:
        a, b, c = _case(rng, m, n, k, batch, ta, tb)
        host, dev = _run_both(hip, L, a, b, c, ta, tb, 0.75, 0.5)
        ref = _reference(a, b, c, ta, tb, 0.75, 0.5)
        print(f"{(m, n, k, batch, ta, tb)}: |allocating - numpy| {np.abs(host - ref).max():.3e}, "
              f"|no-alloc - numpy| {np.abs(dev - ref).max():.3e}", flush=True)
        np.testing.assert_allclose(dev, ref, rtol=0, atol=1e-5 * np.sqrt(k))
        np.testing.assert_array_equal(dev, host)  # same kernels, same plan


def test_noalloc_gemm_rejects_bad_arguments():
    L = _lib()
    assert L.pmc_nn_gemm_noalloc(0, 4, 4, 1, 1.0, None, 4, 0, 0, None, 4, 0, 0, 0.0, None, 4, 0, None) == 1
