"""The oracle's CDF wavelet machinery and adapt_tree.

Pinned against
  * closed forms: the CDF4,4 decomposition low-pass (-1,0,18,-16,-63,144,348,
    144,-63,-16,18,0,-1)/512 and its coarse-extension widths, filter sum rules
    and vanishing moments;
  * perfect reconstruction IWT(FWT(u)) = u on periodic blocks, 2D and 3D;
  * the reference's --coarsen-everywhere outputs
    (TESTING/wavelets/adaptive_CDF{22,42,44,62}/vor_00200.h5, produced from the
    --refine-everywhere result of vor_000020000000.h5): identical block lists
    and fields within rel. L2 1e-14 (the files carry rounding noise of their
    own, about 1 ulp); for CDF20/40/60 the coarsened interiors equal the input
    bit for bit, as the reference's files do.
"""
import numpy as np
import pytest

from helpers import fixture_to_hvy, load_npz
from oracle import pyoracle as po

ALL = ["CDF20", "CDF22", "CDF40", "CDF42", "CDF44", "CDF60", "CDF62"]


def test_cdf44_tables():
    w = po.wavelet("CDF44")
    lo, hd = w.filt("HD")
    assert lo == -6
    assert np.array_equal(hd * 512, [-1, 0, 18, -16, -63, 144, 348, 144, -63, -16, 18, 0, -1])
    glo, gd = w.filt("GD")
    assert glo == -3 and np.array_equal(gd * 16, [1, 0, -9, 16, -9, 0, 1])
    rlo, gr = w.filt("GR")
    assert np.array_equal(gr, hd * (-1.0) ** np.arange(-6, 7))
    assert (w.g, w.order) == (6, 4)
    assert (w.Nscl, w.Nwcl, w.Nreconl, w.Nscr, w.Nwcr, w.Nreconr) == (5, 8, 14, 6, 9, 15)
    # the FD minimum 2*reach raises small widths only (CDF40: Nwc 3 -> 4)
    w40 = po.wavelet("CDF40", fd_max_size=2)
    assert (w40.Nwcl, w40.Nwcr, w40.g_rhs) == (4, 4, 2)


@pytest.mark.parametrize("wav", ALL)
def test_filter_identities(wav):
    w = po.wavelet(wav)
    lo, hd = w.filt("HD")
    glo, gd = w.filt("GD")
    rlo, hr = w.filt("HR")
    assert hd.sum() == 1.0 and hr.sum() == 2.0 and hr[-rlo] == 1.0
    i = np.arange(glo, glo + gd.size, dtype=np.float64)
    for k in range(w.X):  # GD annihilates polynomials of degree < X
        assert abs(np.dot(gd, i ** k)) < 1e-12
    # low-pass HD reproduces degree < Y moments (lifted) / delta (unlifted)
    j = np.arange(lo, lo + hd.size, dtype=np.float64)
    for k in range(1, max(w.Y, 1)):
        assert abs(np.dot(hd, j ** k)) < 1e-12


def _periodic_ghosts(u, g, dim):
    """Fill the ghosts of a single block periodically (axis order z, y, x)."""
    axes = [-1, -2, -3][:dim]
    for ax in axes:
        n = u.shape[ax] - 2 * g
        sl = [slice(None)] * u.ndim
        src_lo = sl.copy(); src_lo[ax] = slice(n, n + g)
        dst_lo = sl.copy(); dst_lo[ax] = slice(0, g)
        src_hi = sl.copy(); src_hi[ax] = slice(g, 2 * g)
        dst_hi = sl.copy(); dst_hi[ax] = slice(n + g, n + 2 * g)
        u[tuple(dst_lo)] = u[tuple(src_lo)]
        u[tuple(dst_hi)] = u[tuple(src_hi)]
    return u


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("wav", ALL)
def test_perfect_reconstruction(wav, dim):
    w = po.wavelet(wav)
    g = w.g
    bs = [16] * dim
    rng = np.random.default_rng(7)
    shape = (2,) + ((bs[0] + 2 * g,) * dim if dim == 3 else (1, bs[0] + 2 * g, bs[0] + 2 * g))
    u = _periodic_ghosts(rng.standard_normal(shape), g, dim)
    wd = po.fwt_block(w, dim, bs, g, u)
    wd = _periodic_ghosts(wd, g, dim)    # spaghetti form is periodic for even Bs
    r = po.iwt_block(w, dim, bs, g, wd, np.zeros_like(wd))
    inner = (slice(None),) + ((slice(g, g + bs[0]),) * 3 if dim == 3 else (0, slice(g, g + bs[0]), slice(g, g + bs[0])))
    assert np.max(np.abs(r[inner] - u[inner])) < 1e-12
    # a planted mistake (one HR tap flipped) breaks it
    if wav == "CDF44" and dim == 2:
        bad = po.wavelet(wav)
        bad.HR[0] = -bad.HR[0]
        r2 = po.iwt_block(bad, dim, bs, g, wd, np.zeros_like(wd))
        assert np.max(np.abs(r2[inner] - u[inner])) > 1e-3


@pytest.mark.parametrize("wav", ["CDF20", "CDF40", "CDF60"])
def test_fwt_polynomial_has_no_details(wav):
    w = po.wavelet(wav)
    g, bs = w.g, [16, 16]
    n = bs[0] + 2 * g
    x = np.arange(n, dtype=np.float64)
    X, Y = np.meshgrid(x, x, indexing="xy")
    deg = w.X - 1
    u = ((0.01 * X + 0.2) ** deg + (0.02 * Y - 0.1) ** deg)[None, None]
    wd = po.fwt_block(w, 2, bs, g, u)
    wc = wd[0, 0, g:g + bs[1], g:g + bs[0]].copy()
    wc[::2, ::2] = 0.0
    assert np.max(np.abs(wc)) < 1e-12


def coarsen_everywhere(wav, use_ce=None):
    """wabbit-post --refine-everywhere then --coarsen-everywhere on the
    reference's input file (sparse_to_dense.f90:102-151: Jmin = 1, Jmax from
    the file, coarse extension iff the wavelet is lifted)."""
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
    t2 = po.Tree(w, 2, 1, jmax, [bs, bs], g, tc, lv, hv, use_ce=use_ce)
    t2.sync()
    t2.adapt()
    return (g, bs) + t2.get()


@pytest.mark.parametrize("wav", ALL)
def test_coarsen_everywhere_matches_reference(wav):
    g, bs, tc, lv, hv = coarsen_everywhere(wav)
    src = load_npz("wavelet_input.npz")
    ref = src if wav in ("CDF20", "CDF40", "CDF60") else load_npz(f"wavelet_coarsen_{wav}.npz")
    mine_keys = sorted(zip(lv.tolist(), tc.tolist()))
    ref_keys = sorted(zip(ref["level"].astype(int).tolist(), ref["treecode"].astype(np.int64).tolist()))
    assert mine_keys == ref_keys                      # block list bit-exact
    index = {(int(t), int(l)): n for n, (t, l) in enumerate(zip(tc, lv))}
    mine = np.stack([hv[index[(int(t), int(l))], 0, 0, g:g + bs + 1, g:g + bs + 1]
                     for t, l in zip(ref["treecode"], ref["level"])])
    want = ref["field_vor"]
    inner = (slice(None), slice(0, bs), slice(0, bs))
    if wav in ("CDF20", "CDF40", "CDF60"):
        assert np.array_equal(mine[inner], want[inner])
    else:
        err = np.linalg.norm(mine - want) / np.linalg.norm(want)
        assert err <= 1e-14, err
        assert np.max(np.abs(mine - want)) <= 1e-13


def test_coarse_extension_is_needed():
    """Without coarse extension a lifted wavelet misses the reference by far
    more than rounding (the test above has teeth): CDF22 and CDF62
    depend on it near level jumps, CDF42/44 do not on this input."""
    g, bs, tc, lv, hv = coarsen_everywhere("CDF22", use_ce=0)
    ref = load_npz("wavelet_coarsen_CDF22.npz")
    index = {(int(t), int(l)): n for n, (t, l) in enumerate(zip(tc, lv))}
    mine = np.stack([hv[index[(int(t), int(l))], 0, 0, g:g + bs, g:g + bs]
                     for t, l in zip(ref["treecode"], ref["level"])])
    assert np.max(np.abs(mine - ref["field_vor"][:, :bs, :bs])) > 1e-6
