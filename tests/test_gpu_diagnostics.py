"""drn_variable_difference, the device form of Patch::computeVariableDifference
(patch.cpp:733-746): max_norm equals the oracle's bit for bit; l2_norm, whose
sum the device accumulates in FP64, equals numpy's FP64 sum of the same
state-type squares and the oracle's state-type sum to its rounding."""
import numpy as np
import pytest

from drnum_amd import configs, hip, solver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_variable_difference_matches_the_oracle(gpu, dtype):
    from oracle import oracle as O

    w = configs.jet(core=8, blocks=2)
    og = O.OracleGrid(w, dtype)
    og.init_state()
    gg = solver.GpuGrid(w, dtype)
    rng = np.random.default_rng(11)
    for p in range(len(w.patches)):
        for f in (0, 2):
            U = og.field(p, f)
            U[:] = (1.0 + 0.3 * rng.standard_normal(U.shape)).astype(dtype)
            gg.ctx.upload(p, f, U.copy())
    for p in (0, 3, 7):
        for v1, v2 in ((0, 0), (2, 4)):
            print("patch", p, "vars", v1, v2, np.dtype(dtype).name, flush=True)
            mx, l2 = gg.ctx.variable_difference(p, 0, v1, 2, v2)
            omx, ol2 = og.variable_difference(p, 0, v1, 2, v2)
            assert mx == omx
            d = (og.field(p, 2)[v2] - og.field(p, 0)[v1]).astype(dtype)
            sq = (d * d).astype(dtype).astype(np.float64)
            assert l2 == pytest.approx(float(np.sqrt(sq.sum())), rel=1e-14)
            assert l2 == pytest.approx(ol2, rel=1e-4 if dtype == np.float32 else 1e-12)
    with pytest.raises(RuntimeError, match="variable index"):
        gg.ctx.variable_difference(0, 0, 5, 2, 0)
    gg.close()
