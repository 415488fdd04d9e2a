"""The oracle's adaptive ACM main loop against the reference's
TESTING/acm/acm_CDF44 run (penalised cylinder, p-norm sponge, CDF44,
eps = 1e-3 normalised, mask thresholding, max-level dealiasing, security
zone): starting from the saved t = 0 state, 40 steps to t = 0.05 and 82 to
t = 0.1 -- the reference's iteration counts -- with

  * the block lists (treecodes, levels) identical at both output times,
  * the mask chi identical bit for bit on every saved grid,
  * ux, uy, p within relative L2 1e-12 (observed ~2e-14: the reference's
    files carry rounding noise of their own, see test_oracle_wavelet.py).

This pins the non-skew FD4 RHS with penalisation and sponge, the block dt
with the penalisation and sponge limits, refinement, the threshold indicator
with its normalisation, the mask criterion and the security zone."""
import numpy as np

import acm_cyl_case as A
from helpers import load_npz

G, BS = A.G, A.BS


def _by_key(tc, lv):
    return {(int(a), int(b)): k for k, (a, b) in enumerate(zip(tc, lv))}


def test_mask_on_saved_grids():
    for name in ("t0", "t005", "t010"):
        ref = load_npz(f"acm_cyl_{name}.npz")
        hvy = np.zeros((len(ref["treecode"]), 3, 1, BS + 2 * G, BS + 2 * G))
        tree = A.oracle_tree(ref["treecode"], ref["level"], hvy, int(ref["max_level"]))
        tc, lv, _ = tree.get()
        mk = tree.get_mask(A.oracle_mask_params())
        idx = _by_key(tc, lv)
        mine = np.stack([mk[idx[(int(a), int(b))], 0, 0, G:G + BS + 1, G:G + BS + 1]
                         for a, b in zip(ref["treecode"], ref["level"])])
        assert np.array_equal(mine, ref["field_mask"]), name
        assert (ref["field_mask"] > 0).sum() > 100


def _compare(tree, ref, it, it_ref):
    assert it == it_ref
    tc, lv, hv = tree.get()
    assert sorted(_by_key(tc, lv)) == sorted(_by_key(ref["treecode"], ref["level"]))   # grid bit-exact
    idx = _by_key(tc, lv)
    for c, name in enumerate(["ux", "uy", "p"]):
        mine = np.stack([hv[idx[(int(a), int(b))], c, 0, G:G + BS, G:G + BS]
                         for a, b in zip(ref["treecode"], ref["level"])])
        want = ref["field_" + name][:, :BS, :BS]
        assert np.linalg.norm(mine - want) / np.linalg.norm(want) <= 1e-12, name


def test_adaptive_cylinder_run_matches_reference():
    seen = {}
    ref05 = load_npz("acm_cyl_t005.npz")

    def on_step(tree, time, it):
        if abs(time - 0.05) < 1e-12:
            _compare(tree, ref05, it, int(ref05["iteration"]))
            seen["t005"] = it

    tree, time, it = A.oracle_run(0.1, on_step)
    assert "t005" in seen and abs(time - 0.1) < 1e-12
    ref10 = load_npz("acm_cyl_t010.npz")
    _compare(tree, ref10, it, int(ref10["iteration"]))
