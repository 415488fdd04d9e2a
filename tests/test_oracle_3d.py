"""The oracle's 3D mesh operations, which no 3D adaptive reference fixture
covers (the 3D fixtures, TESTING/acm/taylorGreen, are equidistant), pinned by
closed forms and the invariants of the reference's own unit tests:

  * prediction (module_wavelets.f90:96-284) in 3D reproduces polynomials of
    degree < order exactly, and not one degree above;
  * refineBlock (refinementExecute.f90:4-116) in 3D: every daughter holds the
    mother's cubic at its octant (digit k: x-half mod(k/2,2), y-half
    mod(k,2), z-half k/4, refinementExecute.f90:98-105);
  * sync_ghosts_tree with level jumps in 3D converges with the predictor's
    order: the ghost-node error of cos(2^f 2 pi x)cos(..y)cos(..z) on a
    random graded forest falls by >= 2^(p - 1/2) per halving of the
    frequency, the criterion of unit_test_ghostSync.f90:223-251;
  * two independent restatements of the 3D ghost fill (point location,
    oracle/adapt.c; the tree sync adapt_tree uses, oracle/tree.c) agree bit
    for bit for the unlifted wavelets;
  * coarsen(refine(u)) = u (unit_test_refineCoarsen.f90:129, there on norms,
    here point by point) and adapt(adapt(u)) = adapt(u)
    (unit_test_waveletDecomposition_invertibility.f90:116) in 3D, on both
    branches of wavelet_decompose_full_tree, which agree bit for bit.

A planted mistake (z pass of the 3D prediction of order 2) fails the
convergence check.
"""
import numpy as np
import pytest

from helpers import random_graded_forest
from oracle import pyoracle as po
from wabbit_amd.forest import equidistant_treecodes, ijk_from_treecode


def _poly3(X, Y, Z, deg):
    return (0.3 * X + 0.1) ** deg + (0.2 * Y - 0.4) ** deg + (0.25 * Z + 0.3) ** deg + 0.05 * X * Y * Z


@pytest.mark.parametrize("order", [2, 4, 6])
def test_prediction_3d_exact_on_polynomials(order):
    n = 10
    x = np.arange(n, dtype=np.float64)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    xf = np.arange(2 * n - 1) / 2.0
    ZF, YF, XF = np.meshgrid(xf, xf, xf, indexing="ij")
    m = order - 1
    inner = (slice(m, -m),) * 3
    for deg, exact in ((order - 1, True), (order, False)):
        coarse = np.ascontiguousarray(_poly3(X, Y, Z, deg))
        fine = np.zeros((2 * n - 1,) * 3)
        po.lib().or_prediction(po._d(coarse), n, n, n, po._d(fine), order)
        err = np.max(np.abs(fine[inner] - _poly3(XF, YF, ZF, deg)[inner]))
        assert (err <= 1e-11) if exact else (err > 1e-6), (deg, err)


def test_refine_block_3d_octants():
    """refineBlock of a mother holding a cubic: every daughter's interior is
    the cubic sampled at the daughter's own points."""
    bs, g, order = 8, 3, 4
    n1 = bs + 2 * g
    dxm = 1.0
    ar = (np.arange(n1) - g) * dxm
    Z, Y, X = np.meshgrid(ar, ar, ar, indexing="ij")
    mother = np.ascontiguousarray(_poly3(X, Y, Z, 3)[None])
    nd = 8
    daughters = [np.zeros_like(mother) for _ in range(nd)]
    arr = (po._dp * nd)(*[po._d(d) for d in daughters])
    po.lib().or_refine_block(3, po._i(np.array([bs, bs, bs], np.int32)), g, 1, po._d(mother), arr, order)
    h = dxm / 2
    for k, d in enumerate(daughters):
        off = [((k // 2) % 2) * bs * h, (k % 2) * bs * h, (k // 4) * bs * h]    # x, y, z halves
        ax = [off[i] + (np.arange(n1) - g) * h for i in range(3)]
        ZD, YD, XD = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        sl = (slice(g, g + bs),) * 3
        assert np.max(np.abs(d[0][sl] - _poly3(XD, YD, ZD, 3)[sl])) <= 1e-11, k


def _cos_field(tc, lv, jmax, bs, g, freq):
    n1 = bs + 2 * g
    out = np.zeros((len(tc), 1, n1, n1, n1))
    for k, (t, lev) in enumerate(zip(tc, lv)):
        ijk = ijk_from_treecode(int(t), int(lev), jmax, 3)
        dx = 2.0 ** (-int(lev)) / bs
        xs = [ijk[d] * bs * dx + (np.arange(n1) - g) * dx for d in range(3)]
        Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
        out[k, 0] = np.cos(freq * 2 * np.pi * X) * np.cos(freq * 2 * np.pi * Y) * np.cos(freq * 2 * np.pi * Z)
    return out


def _ghost_errors(wav, plant=0, bs=8):
    w = po.wavelet(wav)
    g = w.g
    jmax = 4
    tc, lv = random_graded_forest(3, 1, jmax, seed=3, nref=12)
    assert len(set(lv.tolist())) == 3
    errs = []
    for f in (1, 2, 4):
        ex = _cos_field(tc, lv, jmax, bs, g, f)
        hv = np.full_like(ex, 9.0e9)
        sl = (slice(None), slice(None), slice(g, g + bs), slice(g, g + bs), slice(g, g + bs))
        hv[sl] = ex[sl]
        t = po.Tree(w, 3, 1, jmax, [bs] * 3, g, tc, lv, hv, plant=plant)
        t.sync()
        tco, lvo, out = t.get()
        t.close()
        key = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(tc, lv))}
        exo = ex[[key[(int(a), int(b))] for a, b in zip(tco, lvo)]]
        errs.append(np.sqrt(np.sum((out - exo) ** 2) / np.sum(exo ** 2)))
    return np.array(errs)


@pytest.mark.parametrize("wav,p", [("CDF20", 2), ("CDF22", 2), ("CDF40", 4), ("CDF44", 4), ("CDF60", 6)])
def test_sync_3d_level_jumps_convergence_order(wav, p):
    # the reference requires g < (Bs + 1) / 2 on adaptive grids (module_mpi.f90:150-163)
    e = _ghost_errors(wav, bs=12 if po.wavelet(wav).g >= 4 else 8)
    ratio = e[1:] / e[:-1]
    assert np.all(np.log2(ratio) >= p - 0.5), np.log2(ratio)
    assert np.mean(np.sqrt(ratio)) >= p - 0.5          # unit_test_ghostSync.f90's own criterion


def test_sync_3d_planted_mistake_fails():
    e = _ghost_errors("CDF40", plant=4)
    assert np.min(np.log2(e[1:] / e[:-1])) < 3.5


@pytest.mark.parametrize("wav", ["CDF20", "CDF40", "CDF60"])
def test_sync_3d_two_restatements_agree(wav):
    w = po.wavelet(wav)
    g, bs, jmax = w.g, 12 if wav == "CDF60" else 8, 4
    tc, lv = random_graded_forest(3, 1, jmax, seed=9, nref=10)
    u = np.random.default_rng(4).standard_normal((len(tc), 2, bs + 2 * g, bs + 2 * g, bs + 2 * g))
    t = po.Tree(w, 3, 1, jmax, [bs] * 3, g, tc, lv, u)
    t.sync()
    tco, lvo, a = t.get()
    t.close()
    key = {(int(x), int(y)): k for k, (x, y) in enumerate(zip(tc, lv))}
    b = np.ascontiguousarray(u[[key[(int(x), int(y))] for x, y in zip(tco, lvo)]])
    f = po.LeafForest(3, jmax, [bs] * 3, g, tco, lvo)
    f.sync(b, g, g, w.order)
    assert np.array_equal(a, b)


def _smooth3(tc, lv, jmax, bs, g, seed=1):
    n1 = bs + 2 * g
    rng = np.random.default_rng(seed)
    modes = [(rng.integers(1, 3, 3), rng.uniform(0, 2 * np.pi), rng.uniform(0.3, 1)) for _ in range(4)]
    out = np.zeros((len(tc), 2, n1, n1, n1))
    for k, (t, lev) in enumerate(zip(tc, lv)):
        ijk = ijk_from_treecode(int(t), int(lev), jmax, 3)
        dx = 2.0 ** (-int(lev)) / bs
        xs = [ijk[d] * bs * dx + (np.arange(n1) - g) * dx for d in range(3)]
        Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
        for c in range(2):
            out[k, c] = sum(a * np.cos(2 * np.pi * (kk[0] * X + kk[1] * Y + kk[2] * Z) + ph + c)
                            for kk, ph, a in modes)
    return out


@pytest.mark.parametrize("wav", ["CDF40", "CDF44", "CDF22"])
@pytest.mark.parametrize("branch", [1, 2])
def test_coarsen_refine_identity_3d(wav, branch):
    w = po.wavelet(wav)
    g, bs, jmax = w.g, 16, 3
    tc, lv = equidistant_treecodes(1, 3, jmax)
    u = _smooth3(tc, lv, jmax, bs, g)
    t = po.Tree(w, 3, 1, jmax, [bs] * 3, g, tc, lv, u, decompose_branch=branch)
    t.sync()
    tc0, lv0, u0 = t.get()
    t.refine_everywhere()
    assert t.size() == 64
    t.adapt()                       # indicator "everywhere": back to Jmin
    tc1, lv1, u1 = t.get()
    assert np.array_equal(tc0, tc1) and np.array_equal(lv0, lv1)
    sl = (slice(None), slice(None), slice(g, g + bs), slice(g, g + bs), slice(g, g + bs))
    assert np.max(np.abs(u1[sl] - u0[sl])) <= 1e-14 * np.max(np.abs(u0[sl])) * 10


def _bump3(tc, lv, jmax, bs, g):
    n1 = bs + 2 * g
    out = np.zeros((len(tc), 2, n1, n1, n1))
    for k, (t, lev) in enumerate(zip(tc, lv)):
        ijk = ijk_from_treecode(int(t), int(lev), jmax, 3)
        dx = 2.0 ** (-int(lev)) / bs
        xs = [ijk[d] * bs * dx + (np.arange(n1) - g) * dx for d in range(3)]
        Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")

        def per(a, c):
            d = a - c
            return d - np.round(d)
        r2 = per(X, 0.3) ** 2 + per(Y, 0.6) ** 2 + per(Z, 0.4) ** 2
        out[k, 0] = np.exp(-r2 / (2 * 0.06 ** 2))
        out[k, 1] = np.sin(2 * np.pi * X) * np.exp(-r2 / (2 * 0.1 ** 2))
    return out


def test_adapt_idempotent_3d_both_branches():
    """Bs = 16 with CDF44 is below the leaf-first bound (18): the reference
    takes the level-wise branch here.  Thresholded adaptation to a mixed
    grid; a second adapt_tree changes nothing; both branches agree."""
    w = po.wavelet("CDF44")
    g, bs, jmax = w.g, 16, 3
    tc, lv = random_graded_forest(3, 1, jmax, seed=5, nref=6)
    u = _bump3(tc, lv, jmax, bs, g)
    res = {}
    sl = (slice(None), slice(None), slice(g, g + bs), slice(g, g + bs), slice(g, g + bs))
    for branch in (1, 2):
        t = po.Tree(w, 3, 1, jmax, [bs] * 3, g, tc, lv, u, indicator=1, eps=3e-3, groups=[1, 1],
                    decompose_branch=branch)
        t.sync()
        t.adapt()
        a1 = t.get()
        t.sync()
        t.adapt()
        a2 = t.get()
        t.close()
        assert len(set(a1[1].tolist())) == 2                        # a mixed-level grid
        assert np.array_equal(a1[0], a2[0]) and np.array_equal(a1[1], a2[1])
        assert np.max(np.abs(a1[2][sl] - a2[2][sl])) <= 1e-14
        res[branch] = a1
    assert all(np.array_equal(x, y) for x, y in zip(res[1], res[2]))
