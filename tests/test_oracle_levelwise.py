"""The level-wise branch of wavelet_decompose_full_tree against the
reference's own outputs.

adapt_tree.f90:321-322 takes the leaf-first branch iff every Bs >= 3 max|HD
tap| (CDF44: Bs >= 18); below that (the c2 channel's Bs = 16, the small 3D
grids) it decomposes level by level, finest first, syncing from medium and
finer neighbours only once coarse extension is on (adapt_tree.f90:393-409,
:477-487).  Every reference fixture has Bs >= 26, so all of them ran the
leaf-first branch.  Both branches must give the same decomposition (coarse
extension cuts the dependence on coarser neighbours), so forcing the oracle
onto the level-wise branch on those same inputs must reproduce the
reference's outputs:

  * TESTING/wavelets/adaptive_CDF{20,22,40,42,44,60,62} --coarsen-everywhere
    (Bs = 32): block lists identical, fields within the reference files'
    rounding noise;
  * TESTING/acm/acm_CDF44 (Bs = 26, thresholded adaptation with the mask
    criterion, 40 RK4 steps to t = 0.05): block list and iteration count
    identical, fields within relative L2 1e-12.

Planted mistakes in the level-wise branch (or_adapt_params.plant: coarser
leaves never decomposed; the per-level ghost sync skipped) fail.
"""
import numpy as np
import pytest

import acm_cyl_case as A
from helpers import fixture_to_hvy, load_npz
from oracle import pyoracle as po
from wabbit_amd.params import butcher_rk4, clamp_dt

ALL = ["CDF20", "CDF22", "CDF40", "CDF42", "CDF44", "CDF60", "CDF62"]
LEVEL_WISE, LEAF_FIRST = 2, 1


def coarsen_everywhere(wav, branch, plant=0):
    w = po.wavelet(wav)
    g = w.g
    src = load_npz("wavelet_input.npz")
    bs = int(src["bs"][0])
    jmax = int(src["max_level"])
    hvy = fixture_to_hvy([src["field_vor"]], src["origin"], src["spacing"], bs, g, 2)
    t = po.Tree(w, 2, 1, jmax, [bs, bs], g, src["treecode"], src["level"], hvy)
    t.sync()
    t.refine_everywhere()
    tc, lv, hv = t.get()
    t2 = po.Tree(w, 2, 1, jmax, [bs, bs], g, tc, lv, hv, decompose_branch=branch, plant=plant)
    t2.sync()
    t2.adapt()
    return (g, bs) + t2.get()


def _vs_reference(wav, g, bs, tc, lv, hv):
    src = load_npz("wavelet_input.npz")
    ref = src if wav in ("CDF20", "CDF40", "CDF60") else load_npz(f"wavelet_coarsen_{wav}.npz")
    keys = sorted(zip(lv.tolist(), tc.tolist()))
    ref_keys = sorted(zip(ref["level"].astype(int).tolist(), ref["treecode"].astype(np.int64).tolist()))
    assert keys == ref_keys
    index = {(int(t), int(l)): n for n, (t, l) in enumerate(zip(tc, lv))}
    mine = np.stack([hv[index[(int(t), int(l))], 0, 0, g:g + bs, g:g + bs]
                     for t, l in zip(ref["treecode"], ref["level"])])
    return np.max(np.abs(mine - ref["field_vor"][:, :bs, :bs]))


@pytest.mark.parametrize("wav", ALL)
def test_level_wise_coarsen_everywhere_matches_reference(wav):
    g, bs, tc, lv, hv = coarsen_everywhere(wav, LEVEL_WISE)
    err = _vs_reference(wav, g, bs, tc, lv, hv)
    assert err <= (0.0 if wav in ("CDF20", "CDF40", "CDF60") else 1e-14), err
    # and the two branches agree bit for bit
    _, _, tc1, lv1, hv1 = coarsen_everywhere(wav, LEAF_FIRST)
    assert np.array_equal(tc1, tc) and np.array_equal(lv1, lv) and np.array_equal(hv1, hv)


@pytest.mark.parametrize("wav", ["CDF22", "CDF62"])
def test_level_wise_planted_mistake_fails_coarsening(wav):
    g, bs, tc, lv, hv = coarsen_everywhere(wav, LEVEL_WISE, plant=1)
    assert _vs_reference(wav, g, bs, tc, lv, hv) > 1e-6


def _acm_run(branch, t_end, plant=0, max_steps=None):
    tc, lv, hvy, _ = A.initial_state()
    w = po.wavelet("CDF44", fd_max_size=2)
    tree = po.Tree(w, 2, A.JMIN, A.JMAX, [A.BS, A.BS], A.G, tc, lv, hvy, indicator=1, eps=1e-3,
                   groups=[1, 1, 1], force_maxlevel_dealiasing=True, eps_normalized=True, threshold_mask=True,
                   domain=(A.L, A.L, 1.0), mask=A.oracle_mask_params(), decompose_branch=branch, plant=plant)
    ap, mp, tp, bt = po.make_params(A.acm_params()), A.oracle_mask_params(), A.time_params(), butcher_rk4()
    time, it = 0.0, 0
    while time < t_end - 1e-12 and (max_steps is None or it < max_steps):
        tree.sync()
        tree.refine_everywhere()
        dt = clamp_dt(tree.dt(ap), time, tp)
        tree.rk_step(ap, mp, bt, dt, A.GRHS)
        time += dt
        it += 1
        tree.adapt()
    return tree, time, it


def test_level_wise_acm_cdf44_matches_reference():
    tree, time, it = _acm_run(LEVEL_WISE, 0.05)
    ref = load_npz("acm_cyl_t005.npz")
    assert it == int(ref["iteration"]) and abs(time - 0.05) < 1e-12
    tc, lv, hv = tree.get()
    idx = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(tc, lv))}
    assert sorted(idx) == sorted((int(a), int(b)) for a, b in zip(ref["treecode"], ref["level"]))
    G, BS = A.G, A.BS
    for c, name in enumerate(["ux", "uy", "p"]):
        mine = np.stack([hv[idx[(int(a), int(b))], c, 0, G:G + BS, G:G + BS]
                         for a, b in zip(ref["treecode"], ref["level"])])
        want = ref["field_" + name][:, :BS, :BS]
        assert np.linalg.norm(mine - want) / np.linalg.norm(want) <= 1e-12, name


def test_level_wise_planted_mistakes_fail():
    """Each planted mistake fails: skipping the decomposition of coarser
    leaves leaves the correct run within 3 steps; skipping the per-level
    ghost sync gives another grid or iteration count at t = 0.05."""
    good, _, _ = _acm_run(LEVEL_WISE, 0.05, max_steps=3)
    bad, _, _ = _acm_run(LEVEL_WISE, 0.05, plant=1, max_steps=3)
    tg, lg, hg = good.get()
    tb, lb, hb = bad.get()
    if np.array_equal(tg, tb) and np.array_equal(lg, lb):
        G, BS = A.G, A.BS
        sl = (slice(None), slice(None), 0, slice(G, G + BS), slice(G, G + BS))
        assert np.linalg.norm(hb[sl] - hg[sl]) / np.linalg.norm(hg[sl]) > 1e-9
    bad2, _, it = _acm_run(LEVEL_WISE, 0.05, plant=2)
    ref = load_npz("acm_cyl_t005.npz")
    tc, lv, _ = bad2.get()
    same_grid = sorted(zip(tc.tolist(), lv.tolist())) == sorted(
        (int(a), int(b)) for a, b in zip(ref["treecode"], ref["level"]))
    assert it != int(ref["iteration"]) or not same_grid
