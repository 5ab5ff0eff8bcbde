"""Size limits of the device library, refused with an error before anything is
allocated: patch sides up to 32767 cells, and 32-bit byte offsets within one
variable array (2^30 stored cells in FP32, 2^29 in FP64, patches padded to 64)."""
import numpy as np
import pytest

from drnum_amd import configs, hip

pytestmark = pytest.mark.gpu

PAR = np.zeros((6, configs.NPAR))
SLIP = [configs.BC_SLIP] * 6


def test_patch_side_limit(gpu):
    ctx = hip.Context(np.float32)
    with pytest.raises(RuntimeError, match="32767"):
        ctx.add_patch(32768, 4, 4, [0] * 6, (1.0, 1.0, 1.0), SLIP, PAR)
    assert ctx.add_patch(32767, 1, 1, [0] * 6, (1.0, 1.0, 1.0), SLIP, PAR) == 0
    ctx.close()


@pytest.mark.parametrize("dtype,limit", [(np.float32, 1 << 30), (np.float64, 1 << 29)])
def test_offset_limit(gpu, dtype, limit):
    # two patches of 1024 x 1024 x n cells, just above the limit together
    n = limit // (2 * 1024 * 1024) + 1
    ctx = hip.Context(dtype)
    for _ in range(2):
        ctx.add_patch(1024, 1024, n, [0] * 6, (1.0, 1.0, 1.0), SLIP, PAR)
    with pytest.raises(RuntimeError, match="2\\^32 bytes"):
        ctx.layout()
    ctx.close()
