"""drn_core_sums_noalloc, the variant of drn_core_sums that allocates nothing:
its partial sums live in a fixed device array of the library and its result in
device memory the caller owns.  It runs the same kernels as drn_core_sums, so
its sums must equal drn_core_sums bit for bit: two calls in a row, the second on
a smaller grid (fewer partial rows than the first left in the shared array),
in FP64 and FP32, and after an RK step that changed the state."""
import ctypes as C

import numpy as np
import pytest

from drnum_amd import configs, hip, solver

pytestmark = pytest.mark.gpu

_hip = None


def hiprt():
    """The HIP runtime the library is linked against (test-side device buffers only)."""
    global _hip
    if _hip is None:
        hip.lib()
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        _hip = L
    return _hip


class DeviceDoubles:
    def __init__(self, n):
        self.n = n
        self.p = C.c_void_p()
        assert hiprt().hipMalloc(C.byref(self.p), 8 * n) == 0
        assert hiprt().hipMemset(self.p, 0xFF, 8 * n) == 0  # NaN pattern: a value never written shows

    def read(self):
        out = np.empty(self.n)
        assert hiprt().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, 8 * self.n, 2) == 0  # DeviceToHost
        return out

    def free(self):
        hiprt().hipFree(self.p)


def _grid(core, blocks, dtype, seed):
    w = configs.jet(core=core, blocks=blocks)
    gg = solver.GpuGrid(w, dtype)
    rng = np.random.default_rng(seed)
    for p, P in enumerate(w.patches):
        U = (1.0 + 0.5 * rng.random((5, P.cells))).astype(dtype)
        gg.ctx.upload(p, 0, U)
    return gg


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_core_sums_equal_core_sums(gpu, dtype):
    out = DeviceDoubles(5)
    try:
        big = _grid(16, 4, dtype, 1)      # 64 patches of 20^3: 5,120 plane chunks
        small = _grid(8, 2, dtype, 2)     # 8 patches of 10^3 (or 12^3): far fewer chunks
        for name, gg in (("big", big), ("small", small), ("big again", big)):
            print("noalloc sums:", name, np.dtype(dtype).name, flush=True)
            ref = gg.ctx.core_sums(0)
            gg.ctx.core_sums_noalloc(out.p.value, 0)
            gg.ctx.sync()
            got = out.read()
            assert np.array_equal(got, ref), (name, got, ref)
        # after an RK step from the jet's own initial state, on the smaller grid again
        small.upload_initial()
        small.rk_step()
        ref = small.ctx.core_sums(0)
        small.ctx.core_sums_noalloc(out.p.value, 0)
        small.ctx.sync()
        assert np.array_equal(out.read(), ref)
        big.close()
        small.close()
    finally:
        out.free()


def test_noalloc_core_sums_rejects_bad_arguments(gpu):
    gg = _grid(8, 2, np.float64, 3)
    with pytest.raises(RuntimeError):
        gg.ctx.core_sums_noalloc(0, 0)          # no output buffer
    out = DeviceDoubles(5)
    try:
        with pytest.raises(RuntimeError):
            gg.ctx.core_sums_noalloc(out.p.value, 3)   # no such field
    finally:
        out.free()
    gg.close()
