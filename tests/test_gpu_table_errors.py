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


def test_strides_beyond_the_table_limits_are_refused(gpu):
    """A donor's stride (entries per receiver) must lie in 1..64; the reference
    stops above 8 (gpu_patchiterator.h:226-228), this library takes up to 64."""
    for stride in (0, 65):
        ctx = two_patches()
        t = dict(receivers=np.array([0], dtype=np.int64), donors=[(1, 1, stride, 0, 0)],
                 rec_concat=np.array([0], dtype=np.int64), donor_index=np.arange(max(stride, 1), dtype=np.int64),
                 donor_weight=np.ones(max(stride, 1)))
        with pytest.raises(RuntimeError):
            ctx.set_exchange(0, t)
        ctx.close()


def sequential_sum(groups, states):
    """the exchange's sum in FP32, in the order the reference forms it
    (patch.cpp:573-653): per donor patch the weighted entries one after the
    other, then the donor sums one after the other"""
    tot = np.zeros(5, np.float32)
    for p, idx, w in groups:
        acc = np.zeros(5, np.float32)
        for d, ww in zip(idx, w):
            acc = (acc + states[p][:, d] * np.float32(ww)).astype(np.float32)
        tot = (tot + acc).astype(np.float32)
    return tot


def test_wide_and_multi_donor_sums_keep_the_reference_order(gpu):
    """Receivers with 9 entries from one donor patch (beyond the reference's 8)
    and receivers averaging two donor patches with different strides and
    entry counts (the multi-donor kernel, whose receivers are sorted by entry
    count and whose entries are interleaved per wavefront): every receiver
    equals the sequential FP32 sum bit for bit."""
    ctx = hip.Context(np.float32)
    for _ in range(3):
        ctx.add_patch(N, N, N, [0] * 6, (1.0, 1.0, 1.0), SLIP, PAR)
    ctx.layout()
    rng = np.random.default_rng(21)
    # patch 0: receivers 0..69, each with a random number of entries from
    # patch 1 (1..9) and, for the even ones, from patch 2 (1..5)
    recs = list(range(70))
    plan = {}
    for r in recs:
        g = [(1, rng.choice(np.arange(100, 400), size=int(rng.integers(1, 10)), replace=False),
              rng.uniform(0.05, 0.3, size=9))]
        if r % 2 == 0:
            g.append((2, rng.choice(np.arange(100, 400), size=int(rng.integers(1, 6)), replace=False),
                      rng.uniform(0.05, 0.3, size=5)))
        plan[r] = [(p, idx, w[:len(idx)]) for p, idx, w in g]
    # one donor entry per (donor patch, stride): its receivers, their entries in order
    donors, rec_concat, donor_index, donor_weight = [], [], [], []
    for p in (1, 2):
        for stride in range(1, 10):
            rs = [r for r in recs if any(q == p and len(idx) == stride for q, idx, _ in plan[r])]
            if not rs:
                continue
            donors.append((p, len(rs), stride, len(rec_concat), len(donor_index)))
            for r in rs:
                idx, w = next((idx, w) for q, idx, w in plan[r] if q == p)
                rec_concat.append(r)
                donor_index.extend(int(x) for x in idx)
                donor_weight.extend(float(x) for x in w)
    ctx.set_exchange(0, dict(receivers=np.array(recs, dtype=np.int64), donors=donors,
                             rec_concat=np.array(rec_concat, dtype=np.int64),
                             donor_index=np.array(donor_index, dtype=np.int64),
                             donor_weight=np.array(donor_weight)))
    ctx.commit_exchange()
    states = [rng.uniform(0.5, 2.0, size=(5, N * N * N)).astype(np.float32) for _ in range(3)]
    for p in range(3):
        ctx.upload(p, 0, states[p])
    ctx.exchange(0)
    got = ctx.download(0, 0)
    for r in recs:
        # the device sums the donor patches in the order of the donor entries: patch 1, then patch 2
        want = sequential_sum(sorted(plan[r], key=lambda g: g[0]), states)
        np.testing.assert_array_equal(got[:, r], want, err_msg="receiver %d" % r)
    np.testing.assert_array_equal(got[:, 70:], states[0][:, 70:])
    ctx.close()
