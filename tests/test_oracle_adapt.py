"""The oracle's non-uniform ghost fill and refinement against the reference's
--refine-everywhere outputs (TESTING/wavelets/adaptive_CDF{20,40,60}/vor_00100.h5:
sync_ghosts_tree on a 112-block grid with levels 2-5, refineBlock of every
block, sync before saving), and prediction against polynomials."""
import numpy as np
import pytest

from helpers import fixture_to_hvy, load_npz, params_of
from oracle import pyoracle as po

WAV = {"CDF20": (2, 1), "CDF40": (4, 3), "CDF60": (6, 5)}  # order, g = X-1


def refine_everywhere(wav):
    order, g = WAV[wav]
    src = load_npz("wavelet_input.npz")
    bs = int(src["bs"][0])
    jmax = int(src["max_level"])
    hvy = fixture_to_hvy([src["field_vor"]], src["origin"], src["spacing"], bs, g, 2)
    f = po.LeafForest(2, jmax, [bs, bs], g, src["treecode"], src["level"])
    f.sync(hvy, g, g, order)
    tc, lev, hv2 = f.refine_all(hvy, order)
    f2 = po.LeafForest(2, jmax, [bs, bs], g, tc, lev)
    f2.sync(hv2, g, g, order)
    return tc, lev, hv2, g, bs


@pytest.mark.parametrize("wav", ["CDF20", "CDF40", "CDF60"])
def test_refine_everywhere_matches_reference(wav):
    tc, lev, hv2, g, bs = refine_everywhere(wav)
    ref = load_npz(f"wavelet_refine_{wav}.npz")
    prm = params_of(ref)
    assert len(tc) == (prm.get("n_total") or ref["field_vor"].shape[0])
    index = {(int(t), int(l)): n for n, (t, l) in enumerate(zip(tc, lev))}
    mine = np.stack([hv2[index[(int(t), int(l))], 0, 0, g:g + bs + 1, g:g + bs + 1]
                     for t, l in zip(ref["treecode"], ref["level"])])
    want = ref["field_vor"]
    inner = (slice(None), slice(0, bs), slice(0, bs))
    err = np.linalg.norm(mine[inner] - want[inner]) / np.linalg.norm(want[inner])
    assert err <= 1e-14, err
    assert np.array_equal(mine[inner], want[inner])     # interiors bit-identical
    assert np.array_equal(mine, want)                    # and the synced extra point


@pytest.mark.parametrize("order", [2, 4, 6])
def test_prediction_exact_on_polynomials(order):
    """Interpolating prediction of order p reproduces polynomials of degree < p
    exactly away from the array edges (the closed form the stencils satisfy)."""
    n = 12
    x = np.arange(n, dtype=np.float64)
    X, Y = np.meshgrid(x, x, indexing="xy")
    deg = order - 1
    coarse = np.ascontiguousarray((0.3 * X + 0.1) ** deg + (0.2 * Y - 0.4) ** deg + 0.05 * X * Y)
    fine = np.zeros((2 * n - 1, 2 * n - 1))
    po.lib().or_prediction(po._d(coarse), n, n, 1, po._d(fine), order)
    xf = np.arange(2 * n - 1) / 2.0
    XF, YF = np.meshgrid(xf, xf, indexing="xy")
    exact = (0.3 * XF + 0.1) ** deg + (0.2 * YF - 0.4) ** deg + 0.05 * XF * YF
    m = order - 1
    assert np.allclose(fine[m:-m, m:-m], exact[m:-m, m:-m], rtol=1e-12, atol=1e-12)
    # a planted mistake (wrong sign of the outer coefficients) is caught
    if order >= 4:
        bad = exact.copy()
        bad[1::2, :] += 1e-3
        assert not np.allclose(fine[m:-m, m:-m], bad[m:-m, m:-m], rtol=1e-12, atol=1e-12)
