"""Exchange tables the device library refuses with an error before any kernel
runs: a donor cell that is also a receiver cell (the device fills all
receivers at once, the reference patch after patch, so the result would depend
on the order), cells outside their patch, and null table pointers.  Tables
that pass are applied exactly (unit weights copy the donor values)."""
import numpy as np
import pytest

from drnum_amd import configs, hip

pytestmark = pytest.mark.gpu

PAR = np.zeros((6, configs.NPAR))
SLIP = [configs.BC_SLIP] * 6
N = 8


def two_patches():
    ctx = hip.Context(np.float32)
    for _ in range(2):
        ctx.add_patch(N, N, N, [0] * 6, (1.0, 1.0, 1.0), SLIP, PAR)
    ctx.layout()
    return ctx


def unit_table(rec, don, donor_patch):
    """receivers rec of one patch, each copying donor cell don[i] of donor_patch"""
    return dict(receivers=np.array(sorted(set(rec)), dtype=np.int64), donors=[(donor_patch, len(rec), 1, 0, 0)],
                rec_concat=np.array(rec, dtype=np.int64), donor_index=np.array(don, dtype=np.int64),
                donor_weight=np.ones(len(rec)))


def test_donor_that_is_a_receiver_is_refused(gpu):
    ctx = two_patches()
    ctx.set_exchange(0, unit_table([0, 1, 2, 3], [10, 11, 12, 13], 1))
    ctx.set_exchange(1, unit_table([10], [100], 0))   # cell 10 of patch 1 donates to patch 0
    with pytest.raises(RuntimeError, match="donor cell is also a receiver"):
        ctx.commit_exchange()
    # the same tables without the clash commit, and the exchange copies exactly
    ctx.set_exchange(1, unit_table([20], [100], 0))
    ctx.commit_exchange()
    rng = np.random.default_rng(5)
    states = [rng.uniform(0.5, 2.0, size=(5, N * N * N)).astype(np.float32) for _ in range(2)]
    for p in range(2):
        ctx.upload(p, 0, states[p])
    ctx.exchange(0)
    got = [ctx.download(p, 0) for p in range(2)]
    expect = [s.copy() for s in states]
    expect[0][:, [0, 1, 2, 3]] = states[1][:, [10, 11, 12, 13]]
    expect[1][:, 20] = states[0][:, 100]
    for p in range(2):
        np.testing.assert_array_equal(got[p], expect[p])
    ctx.close()


@pytest.mark.parametrize("rec,don,msg", [([N ** 3], [0], "receiver cell outside patch"),
                                         ([0], [N ** 3], "donor cell outside donor patch"),
                                         ([-1], [0], "receiver cell outside patch")])
def test_cells_outside_their_patch_are_refused(gpu, rec, don, msg):
    ctx = two_patches()
    with pytest.raises(RuntimeError, match=msg):
        ctx.set_exchange(0, unit_table(rec, don, 1))
    ctx.close()


def test_null_table_pointers_are_refused(gpu):
    ctx = two_patches()
    rc = ctx.L.drn_exchange_set(ctx.h, 0, 1, None, None, 0, None, 0, None, 0, None, None)
    assert rc == -1   # DRN_E_ARG
    ctx.close()
