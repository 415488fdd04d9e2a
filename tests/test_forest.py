"""Host forest (C++ wh_forest_*): the ghost-patch tables it builds, applied on
the host, fill every ghost point exactly as the oracle's point-location fill
does (random 2:1-balanced periodic forests, 2D and 3D, predictor orders 2/4/6,
full and RHS ghost depths).  Only host metadata calls: no GPU needed."""
import numpy as np
import pytest

from helpers import apply_patches_host, random_block_data, random_graded_forest
from oracle import pyoracle as po
from wabbit_amd.tree import Forest


CASES = [
    # dim, jmin, jmax, bs, g, order, depth
    (2, 2, 5, 8, 3, 4, 3),
    (2, 2, 5, 8, 3, 4, 2),
    (2, 1, 4, 8, 1, 2, 1),
    (2, 2, 5, 12, 5, 6, 5),
    (3, 1, 3, 6, 3, 4, 3),
    (3, 1, 3, 6, 3, 4, 2),
]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("seed", [1, 2])
def test_patch_tables_match_point_location(case, seed):
    dim, jmin, jmax, bs, g, order, depth = case
    tcs, lev = random_graded_forest(dim, jmin, jmax, seed, nref=6 if dim == 2 else 3)
    n = tcs.shape[0]
    hvy = random_block_data(n, 2, dim, bs, g, seed)
    f = Forest(dim, jmax, [bs] * dim, g, [1.0, 2.0, 1.5][:dim], nb_max=n + 8)
    f.set_blocks(tcs, lev, np.arange(n, dtype=np.int32))
    pt = f.patches(depth, depth, order)
    got = apply_patches_host(hvy, pt, order)
    want = hvy.copy()
    po.LeafForest(dim, jmax, [bs] * dim, g, tcs, lev).sync(want, depth, depth, order)
    assert np.array_equal(got, want)
    assert set(np.unique(pt["op"])) <= {0, 1, 2}
    if len(np.unique(lev)) > 1:
        assert (pt["op"] == 1).any() and (pt["op"] == 2).any()
    f.close()


def test_refine_bookkeeping():
    f = Forest(2, 4, [8, 8], 3, [1.0, 1.0], nb_max=40)
    tcs, lev = random_graded_forest(2, 1, 4, 0, 0)
    f.set_blocks(tcs, lev, np.arange(4, dtype=np.int32))
    tri = f.refine(np.array([0, 2], np.int32))
    assert tri.shape == (8, 3)
    tc, lv, hv, x = f.blocks()
    assert f.n == 4 - 2 + 8 and (lv == 2).sum() == 8
    assert len(set(hv.tolist())) == f.n
    assert not set(tri[:, 1].tolist()) & {0, 2}
    f.close()


def test_slot_base_and_errors():
    """A forest restricted to a slot range of a shared pool: ids outside the
    range, duplicates and levels above Jmax are rejected; refinement takes
    only slots of the range and reports running out of them."""
    from wabbit_amd import lib as L
    f = Forest(2, 4, [8, 8], 3, [1.0, 1.0], nb_max=8, slot_base=100)
    tcs, lev = random_graded_forest(2, 1, 4, 0, 0)           # 4 blocks on level 1
    with pytest.raises(L.WabbitHipError):
        f.set_blocks(tcs, lev, np.arange(4, dtype=np.int32))      # below the range
    with pytest.raises(L.WabbitHipError):
        f.set_blocks(tcs, lev, np.array([100, 101, 101, 102], np.int32))  # duplicate
    with pytest.raises(L.WabbitHipError):
        f.set_blocks(tcs, np.full(4, 5, np.int32), np.arange(100, 104, dtype=np.int32))  # level > jmax
    f.set_blocks(tcs, lev, np.arange(100, 104, dtype=np.int32))
    tri = f.refine(np.array([100], np.int32))
    assert set(tri[:, 1].tolist()) <= set(range(100, 108))
    with pytest.raises(L.WabbitHipError):
        f.refine(np.array([101, 102], np.int32))                  # 8 more slots needed, 1 free + 1 mother
    # the failed call changed nothing
    assert f.n == 3 + 4
    # the slot base can only change while the forest is empty
    with pytest.raises(L.WabbitHipError):
        L.check(L.load().wh_forest_set_slot_base(f.h, 0))
    f.close()


def test_empty_and_single_block_forests():
    """An empty forest has no patches; a single periodic block is its own
    neighbour in every direction (same-level copies only)."""
    f = Forest(2, 3, [8, 8], 3, [1.0, 1.0], nb_max=4)
    f.set_blocks(np.zeros(0, np.int64), np.zeros(0, np.int32), np.zeros(0, np.int32))
    assert f.n == 0 and len(f.patches(3, 3, 4)) == 0
    f.set_blocks(np.zeros(1, np.int64), np.zeros(1, np.int32), np.zeros(1, np.int32))
    pt = f.patches(3, 3, 4)
    assert len(pt) == 8 and np.all(pt["op"] == 0) and np.all(pt["send"] == 0) and np.all(pt["recv"] == 0)
    f.close()


@pytest.mark.parametrize("level", [1, 12])
def test_duplicate_block_rejected(level):
    """Two blocks with one (level, treecode) are rejected by set_blocks, on a
    level held in the dense per-level lookup tables (1) and on one beyond
    them, held in the hash map (12 in 2D); distinct blocks are accepted.  A
    list with a bad slot id leaves the forest unchanged, one with a duplicate
    leaves it empty."""
    from wabbit_amd import lib as L
    from helpers import _tc
    f = Forest(2, 12, [8, 8], 3, [1.0, 1.0], nb_max=8)
    n = 2 ** level
    cells = [(0, 0), (1, 0), (0, 1), (n - 1, n - 1)]
    tcs = np.array([_tc(c, level, 12, 2) for c in cells], np.int64)
    lev = np.full(4, level, np.int32)
    f.set_blocks(tcs, lev, np.arange(4, dtype=np.int32))
    assert f.n == 4
    with pytest.raises(L.WabbitHipError):                          # bad hvy id: forest unchanged
        f.set_blocks(tcs, lev, np.array([0, 1, 2, 99], np.int32))
    assert f.n == 4 and sorted(f.blocks()[0].tolist()) == sorted(tcs.tolist())
    dup = tcs.copy()
    dup[3] = dup[1]
    with pytest.raises(L.WabbitHipError):                          # duplicate: forest left empty
        f.set_blocks(dup, lev, np.arange(4, dtype=np.int32))
    assert f.n == 0
    f.set_blocks(tcs[:3], lev[:3], np.arange(3, dtype=np.int32))   # usable again after the error
    assert f.n == 3
    f.close()


def test_rejected_refine_leaves_forest_unchanged():
    """A refine list with a bad entry after good ones (a block not active, a
    block listed twice, one at Jmax) is rejected before any slot is taken:
    the same valid refine succeeds afterwards, on the same free slots."""
    from wabbit_amd import lib as L
    f = Forest(2, 2, [8, 8], 3, [1.0, 1.0], nb_max=10)
    tcs, lev = random_graded_forest(2, 1, 2, 0, 0)                # 4 blocks on level 1
    f.set_blocks(tcs[:2], lev[:2], np.arange(2, dtype=np.int32))
    for bad in ([0, 7], [0, 0]):
        with pytest.raises(L.WabbitHipError):
            f.refine(np.array(bad, np.int32))
        assert f.n == 2
    tri = f.refine(np.array([0], np.int32))                        # needs 4 of the 8 free slots
    assert tri.shape == (4, 3) and f.n == 5
    tc, lv, hv, _ = f.blocks()
    with pytest.raises(L.WabbitHipError):                          # level 2 = Jmax, after a good entry
        f.refine(np.array([1, int(hv[lv == 2][0])], np.int32))
    assert f.n == 5
    tri = f.refine(np.array([1], np.int32))                        # 4 free + the first mother's slot
    assert tri.shape == (4, 3) and f.n == 8
    f.close()


def test_treecode_digits_beyond_level_rejected():
    """A treecode with a non-zero digit below its block's level (or above
    jmax) is rejected, and the forest is left as it was."""
    from wabbit_amd import lib as L
    from helpers import _tc
    f = Forest(2, 4, [8, 8], 3, [1.0, 1.0], nb_max=8)
    good = np.array([_tc((1, 0), 1, 4, 2)], np.int64)
    f.set_blocks(good, np.array([1], np.int32), np.zeros(1, np.int32))
    for bad in (good[0] | 1, good[0] | (1 << (4 * 2))):
        with pytest.raises(L.WabbitHipError):
            f.set_blocks(np.array([bad], np.int64), np.array([1], np.int32), np.zeros(1, np.int32))
        assert f.n == 1 and f.blocks()[0][0] == good[0]
    f.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_patch_lists_reused_for_a_repeated_block_list(dim):
    """The patch generation is reused when the block list repeats in
    treecodes and levels (as after a refine/coarsen cycle that lands on the
    same grid) but sits in other pool slots: the lists name the new slots,
    equal to a fresh forest's, and a changed block list is generated anew."""
    bs, g, jmax = (8, 3, 5) if dim == 2 else (6, 3, 3)
    tcs, lev = random_graded_forest(dim, 1, jmax, 4, nref=6 if dim == 2 else 3)
    n = tcs.shape[0]
    nb = 3 * n
    rng = np.random.default_rng(9)
    f = Forest(dim, jmax, [bs] * dim, g, [1.0] * dim, nb_max=nb)
    try:
        f.set_blocks(tcs, lev, np.arange(n, dtype=np.int32))
        first = f.patches(g, g, 4)
        for trial in range(3):
            slots = rng.permutation(nb)[:n].astype(np.int32)
            f.set_blocks(tcs, lev, slots)
            got = f.patches(g, g, 4)
            fresh = Forest(dim, jmax, [bs] * dim, g, [1.0] * dim, nb_max=nb)
            fresh.set_blocks(tcs, lev, slots)
            want = fresh.patches(g, g, 4)
            fresh.close()
            assert np.array_equal(got, want), trial
            assert got.size == first.size
        # a different grid (one block fewer) is not taken for the cached one
        f.set_blocks(tcs[:-1], lev[:-1], np.arange(n - 1, dtype=np.int32))
        fresh = Forest(dim, jmax, [bs] * dim, g, [1.0] * dim, nb_max=nb)
        fresh.set_blocks(tcs[:-1], lev[:-1], np.arange(n - 1, dtype=np.int32))
        assert np.array_equal(f.patches(g, g, 4), fresh.patches(g, g, 4))
        fresh.close()
    finally:
        f.close()
