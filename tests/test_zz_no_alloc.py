"""tf_mlp_grad_noalloc (the MLP forward + backward with its work buffers in
one fixed device array) against tf_mlp_grad (the same kernels, buffers
allocated per call) through the C ABI: outputs y bit for bit, the gradient
within fp32 summation rounding (the two calls sum the per-row terms over
different block partitions). Two calls in a row on one stream, the second
with fewer rows, so that what the first left in the shared array must not
leak into the second."""
import ctypes as C

import numpy as np
import pytest

from tetfit_amd import _ffi

# relative L2 bound on each parameter block's gradient: the per-row terms
# are the same, only the fp32 summation order differs (measured ~1e-7)
GRAD_RTOL = 2e-5

BLOCKS = {"W1": slice(0, 16 * 64), "b1": slice(1024, 1088), "W2": slice(1088, 1088 + 64 * 64),
          "b2": slice(5184, 5248), "w3": slice(5248, 5312), "b3": slice(5312, 5313)}


def bf16_round(a):
    """float32 -> nearest bf16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


class _Hip:
    """Device buffers and a stream from the HIP runtime libtetfit_hip.so is
    linked against (one HIP runtime per process: a second one, such as a
    framework's bundled copy, need not see the GPU once this one holds it)."""

    def __init__(self):
        _ffi.hip_lib()  # loads the runtime as the library's dependency
        H = C.CDLL("libamdhip64.so.7")
        vp = C.c_void_p
        H.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
        H.hipFree.argtypes = [vp]
        H.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
        H.hipStreamCreate.argtypes = [C.POINTER(vp)]
        H.hipStreamSynchronize.argtypes = [vp]
        H.hipStreamDestroy.argtypes = [vp]
        self.H = H
        self.bufs = []

    def check(self, rc, what):
        assert rc == 0, f"{what}: hipError {rc}"

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.check(self.H.hipMalloc(C.byref(p), max(a.nbytes, 4)), "hipMalloc")
        self.bufs.append(p)
        self.check(self.H.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        self.check(self.H.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self):
        for p in self.bufs:
            self.H.hipFree(p)
        self.bufs = []


@pytest.mark.gpu
def test_noalloc_matches_allocating_entry_point():
    from tetfit_amd.device import mlp_grad
    assert _ffi.MLP_PARAMS == 5313
    L = _ffi.hip_lib()
    hip = _Hip()
    rng = np.random.default_rng(11)
    P = (rng.standard_normal(_ffi.MLP_PARAMS) * 0.2).astype(np.float32)
    P[BLOCKS["W1"]] = bf16_round(P[BLOCKS["W1"]])
    P[BLOCKS["W2"]] = bf16_round(P[BLOCKS["W2"]])
    stream = C.c_void_p()
    hip.check(hip.H.hipStreamCreate(C.byref(stream)), "hipStreamCreate")
    try:
        dP = hip.upload(P)
        # first call: more tiles than the partial-gradient slots (capped grid)
        # and a 7-row tail; second call: 2 tiles and a 5-row tail
        for n in (20007, 37):
            print(f"case: tf_mlp_grad_noalloc n={n}", flush=True)
            X = bf16_round(rng.standard_normal((n, 16)).astype(np.float32))
            gy = rng.standard_normal(n).astype(np.float32)
            y_ref, g_ref = mlp_grad(P, X, gy)
            bits = (X.view(np.uint32) >> 16).astype(np.uint16)
            dX, dG = hip.upload(bits), hip.upload(gy)
            dY = hip.upload(np.full(n, np.nan, np.float32))
            dGrad = hip.upload(np.full(_ffi.MLP_PARAMS, np.nan, np.float32))
            rc = L.tf_mlp_grad_noalloc(dP, dX, dG, n, dY, dGrad, stream)
            assert rc == 0, L.tf_last_error().decode()
            hip.check(hip.H.hipStreamSynchronize(stream), "hipStreamSynchronize")
            y = hip.download(dY, y_ref)
            g = hip.download(dGrad, g_ref)
            assert np.array_equal(y.view(np.uint32), y_ref.view(np.uint32)), n
            for name, sl in BLOCKS.items():
                assert rel(g[sl], g_ref[sl]) <= GRAD_RTOL, (n, name, rel(g[sl], g_ref[sl]))
            # the bar catches a dropped tail tile: its rows' own gradient is far
            # above the tolerance
            tail = n % 16
            _, g_tail = mlp_grad(P, X[n - tail:], gy[n - tail:])
            assert rel(g_tail, g_ref) > 100 * GRAD_RTOL
    finally:
        hip.H.hipStreamDestroy(stream)
        hip.free()


@pytest.mark.gpu
def test_noalloc_rejects_oversize_and_null():
    L = _ffi.hip_lib()
    one = C.c_void_p(1)
    rc = L.tf_mlp_grad_noalloc(one, one, one, (1 << 31) + 1, one, one, None)
    assert rc != 0 and b"exceeds the maximum" in L.tf_last_error()
    assert L.tf_mlp_grad_noalloc(None, one, one, 4, one, one, None) != 0
