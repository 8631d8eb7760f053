"""The allocation-free pinv (farcur_pinv_noalloc: device pointers, static
device scratch) against farcur_pinv on the same matrices: two calls in a row,
the second smaller, so the scratch left by the first call must not leak into
the second. The device buffers the test passes come from the same HIP
runtime the library links (libamdhip64, opened by its soname once the
library is loaded); the library call itself allocates nothing."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    def __init__(self, farcur):
        farcur.lib()  # loads libfarcur.so and with it the runtime it links
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.bufs.append(p.value)
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0  # synchronises with the device
        return out

    def fill(self, p, value, shape, dtype):
        a = np.full(shape, value, dtype)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _run_noalloc(hip, ctx, W, atol, rtol):
    r = W.shape[0]
    Wf = np.asfortranarray(W).ravel(order="F")
    dW = hip.put(Wf)
    dU = hip.alloc(r * r * 16)
    ds = hip.alloc(r * 8)
    dk = hip.alloc(4)
    hip.fill(dU, np.nan, r * r, np.complex128)
    hip.fill(ds, np.nan, r, np.float64)
    hip.fill(dk, -1, 1, np.int32)
    ctx.pinv_noalloc(dW, r, atol, rtol, dU, ds, dk)  # on the context's stream
    hip.sync()
    U = hip.get(dU, r * r, np.complex128).reshape((r, r), order="F")
    # the input is left as it was
    assert np.array_equal(hip.get(dW, r * r, np.complex128), Wf)
    return U, hip.get(ds, r, np.float64), int(hip.get(dk, 1, np.int32)[0])


def _matrix(rng, r, graded):
    if not graded:
        return rng.standard_normal((r, r)) + 1j * rng.standard_normal((r, r))
    qa, _ = np.linalg.qr(rng.standard_normal((r, r)) + 1j * rng.standard_normal((r, r)))
    qb, _ = np.linalg.qr(rng.standard_normal((r, r)) + 1j * rng.standard_normal((r, r)))
    sv = 10.0 ** np.linspace(0, -12, r)
    return (qa * sv) @ qb.conj().T


@pytest.mark.parametrize("sizes,graded,atol,rtol", [((200, 40), False, 1e-10, 0.0), ((60, 17), True, 0.0, 1e-6),
                                                     ((256, 100), True, 1e-10, 0.0)])
def test_noalloc_pinv_matches_pinv(farcur, sizes, graded, atol, rtol):
    ctx = farcur.Context(0)
    hip = _Hip(farcur)
    rng = np.random.default_rng(sizes[0] + sizes[1])
    for r in sizes:
        W = _matrix(rng, r, graded)
        U0, s0, k0 = ctx.pinv(W, atol, rtol)
        U1, s1, k1 = _run_noalloc(hip, ctx, W, atol, rtol)
        assert k1 == k0
        # the same kernels and launch shape on the same input: equal to FP64 rounding
        np.testing.assert_allclose(s1, s0, rtol=1e-13, atol=1e-15 * s0[0])
        assert np.linalg.norm(U1 - U0) <= 1e-12 * np.linalg.norm(U0)
    hip.free()
    ctx.close()


def test_noalloc_pinv_rejects_large_r(farcur):
    ctx = farcur.Context(0)
    with pytest.raises(farcur.FarcurError, match="outside"):
        ctx.pinv_noalloc(1, 257, 1e-10, 0.0, 1)
    ctx.close()
