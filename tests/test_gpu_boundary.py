"""GPU: the boundary where the host keeps the refinement flags.

  * wh_fwt_threshold (coarseningIndicator_block -> threshold_block,
    coarseningIndicator_block.f90:12-108, threshold_block.f90:1-130) against
    the oracle's FWT and detail, flags and details bit for bit;
  * wh_adapt_tree_flags (adapt_tree with indicator "nothing (external)",
    coarseningIndicator_tree.f90:55) with host flags against the oracle's
    adapt_tree on the same flags (grid identical, fields bit for bit), and
    all flags -1 against indicator "everywhere";
  * the level-wise branch forced on the device on the reference's
    --coarsen-everywhere input (Bs = 32 would take the leaf-first branch),
    against the reference's output files.
"""
import numpy as np
import pytest

from helpers import fixture_to_hvy, load_npz
from oracle import pyoracle as po
from wabbit_amd import lib as L
from wabbit_amd import wavelets as W
from wabbit_amd.tree import Forest, refine_on_device

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("wav,dim,bs", [("CDF44", 2, 16), ("CDF40", 2, 16), ("CDF44", 3, 12), ("CDF22", 3, 12)])
def test_fwt_threshold_matches_oracle(wav, dim, bs):
    w, wo = W.setup(wav), po.wavelet(wav)
    g, neq, n = w.g, 3, 9
    rng = np.random.default_rng(5 + dim)
    nz = bs + 2 * g if dim == 3 else 1
    # smooth blocks of varying amplitude, so that some pass the threshold
    x = np.arange(bs + 2 * g) / (bs + 2 * g)
    base = np.sin(2 * np.pi * x)[None, :] * np.cos(2 * np.pi * x)[:, None]
    src = np.stack([np.stack([(0.01 * (k + 1)) * base[None] + 1e-3 * (k % 3) * rng.standard_normal((nz,) + base.shape)
                              for _ in range(neq)]) for k in range(n)])
    pool = L.Pool(dim, [bs] * dim, g, neq, n, n_work_slots=2, n_tmp_eqn=neq)
    W.set_wavelet(pool, w)
    pool.upload(L.WH_TMP, 0, src)
    pool.upload(L.WH_BLOCK, 0, src)      # outside the transformed region dst keeps its values
    groups = [1, 1, 0]
    ids = np.array([4, 0, 8, 2, 6, 1, 3, 7, 5], np.int32)
    # a threshold between the smallest and largest details
    det_o = np.stack([po.detail_block(dim, [bs] * dim, g, po.fwt_block(wo, dim, [bs] * dim, g, src[b]), groups)
                      for b in ids])
    thr = np.array([np.median(det_o[:, 0]), np.median(det_o[:, 1]), 0.0])
    flags, det = W.fwt_threshold(pool, L.WH_TMP, L.WH_BLOCK, ids, thr, groups)
    assert np.array_equal(det, det_o)
    want = np.where(np.all(det_o <= thr[None, :], axis=1), -1, 0)
    assert np.array_equal(flags, want)
    assert 0 < (flags == -1).sum() < n
    # the decomposed blocks are the oracle's FWT (SC/WC in spaghetti order)
    got = pool.download(L.WH_BLOCK, 0, n)
    for b in ids:
        assert np.array_equal(got[b], po.fwt_block(wo, dim, [bs] * dim, g, src[b]))
    pool.close()


def _fixture_forest(wav):
    w = W.setup(wav)
    g = w.g
    src = load_npz("wavelet_input.npz")
    bs, jmax = int(src["bs"][0]), int(src["max_level"])
    n = src["treecode"].shape[0]
    hvy = fixture_to_hvy([src["field_vor"]], src["origin"], src["spacing"], bs, g, 2)
    nb_max = 4 * n + 300
    f = Forest(2, jmax, [bs, bs], g, [float(src["domain_size"][0])] * 2, nb_max=nb_max)
    f.set_blocks(src["treecode"], src["level"], np.arange(n, dtype=np.int32))
    pool = L.Pool(2, [bs, bs], g, 1, nb_max, n_work_slots=2, n_tmp_eqn=1)
    W.set_wavelet(pool, w)
    f.attach(pool, g, w.order)
    pool.upload(L.WH_BLOCK, 0, hvy)
    return w, g, bs, jmax, f, pool, src, hvy


def _leaves(f, pool):
    tc, lv, hv, _ = f.blocks()
    order = np.lexsort((tc, lv))
    allb = pool.download(L.WH_BLOCK, 0, pool.nb_max)
    return tc[order], lv[order], allb[hv[order]]


def _family_flags(tc, lv, jmax, dim=2):
    """-1 for the daughters of two thirds of the mothers (chosen by a hash of
    the mother), 0 for the rest: whole families, so that coarsening happens."""
    par = [(int(l), int(t) >> ((jmax - int(l) + 1) * dim)) for t, l in zip(tc, lv)]
    return np.array([-1 if ((p[1] * 2654435761 + p[0]) % 3) else 0 for p in par], np.int32)


@pytest.mark.parametrize("wav", ["CDF44", "CDF40", "CDF22"])
def test_adapt_external_flags_matches_oracle(wav):
    w, g, bs, jmax, f, pool, src, hvy = _fixture_forest(wav)
    W.sync_tree(pool, f)
    tc0, lv0, hv0, _ = f.blocks()
    flags = _family_flags(tc0, lv0, jmax)
    W.adapt_tree_flags(pool, f, W.adapt_params(1, jmax, W.WH_IND_EXTERNAL, use_ce=bool(w.lifted)), flags)
    tc, lv, data = _leaves(f, pool)
    assert len(tc) < len(tc0)
    # the oracle on the same leaves and flags (flags in its (level, treecode) order)
    t = po.Tree(po.wavelet(wav), 2, 1, jmax, [bs, bs], g, src["treecode"], src["level"], hvy, indicator=2)
    t.sync()
    key = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(tc0, lv0))}
    to, lo, _ = t.get()
    t.adapt_flags(flags[[key[(int(a), int(b))] for a, b in zip(to, lo)]])
    tco, lvo, want = t.get()
    assert np.array_equal(tco, tc) and np.array_equal(lvo, lv)
    sl = (slice(None), slice(None), 0, slice(g, g + bs), slice(g, g + bs))
    assert np.array_equal(data[sl], want[sl])
    pool.close()
    f.close()


def test_external_all_coarsen_equals_everywhere():
    res = []
    for ind in (W.WH_IND_EVERYWHERE, W.WH_IND_EXTERNAL):
        w, g, bs, jmax, f, pool, src, hvy = _fixture_forest("CDF44")
        W.sync_tree(pool, f)
        p = W.adapt_params(1, jmax, ind, use_ce=True)
        if ind == W.WH_IND_EXTERNAL:
            W.adapt_tree_flags(pool, f, p, -np.ones(f.n, np.int32))
        else:
            W.adapt_tree(pool, f, p)
        res.append(_leaves(f, pool))
        pool.close()
        f.close()
    assert all(np.array_equal(a, b) for a, b in zip(res[0], res[1]))


def test_external_indicator_rejected_without_flags():
    w, g, bs, jmax, f, pool, src, hvy = _fixture_forest("CDF44")
    with pytest.raises(L.WabbitHipError):
        W.adapt_tree(pool, f, W.adapt_params(1, jmax, W.WH_IND_EXTERNAL, use_ce=True))
    with pytest.raises(L.WabbitHipError):
        W.adapt_tree_flags(pool, f, W.adapt_params(1, jmax, W.WH_IND_EXTERNAL, use_ce=True),
                           np.full(f.n, 2, np.int32))
    pool.close()
    f.close()


@pytest.mark.parametrize("wav", ["CDF22", "CDF44", "CDF62", "CDF40"])
def test_level_wise_branch_coarsen_everywhere_reference(wav):
    """--refine-everywhere then --coarsen-everywhere on the device with the
    level-wise decomposition forced: the reference's output files."""
    w, g, bs, jmax, f, pool, src, hvy = _fixture_forest(wav)
    n = src["treecode"].shape[0]
    refine_on_device(pool, f, np.arange(n, dtype=np.int32), w.order, g, presync="tree")
    W.sync_tree(pool, f)
    W.adapt_tree(pool, f, W.adapt_params(1, jmax, W.WH_IND_EVERYWHERE, use_ce=bool(w.lifted), decompose_branch=2))
    tc, lv, data = _leaves(f, pool)
    ref = src if wav == "CDF40" else load_npz(f"wavelet_coarsen_{wav}.npz")
    assert sorted(zip(lv.tolist(), tc.tolist())) == sorted(
        zip(ref["level"].astype(int).tolist(), ref["treecode"].astype(np.int64).tolist()))
    index = {(int(t), int(l)): k for k, (t, l) in enumerate(zip(tc, lv))}
    mine = np.stack([data[index[(int(t), int(l))], 0, 0, g:g + bs, g:g + bs]
                     for t, l in zip(ref["treecode"], ref["level"])])
    err = np.max(np.abs(mine - ref["field_vor"][:, :bs, :bs]))
    assert err <= (0.0 if wav == "CDF40" else 1e-14), err
    pool.close()
    f.close()
