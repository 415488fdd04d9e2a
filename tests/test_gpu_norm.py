"""GPU: eps_norm L1 / L2 / H1 through the C ABI, against the oracle.

threshold_block's renormed details (wh_detail_blocks_norm: the device
reduces the max of |v| per point type, the host applies
wavelet_renorm_block's factors; the rounding is monotone, so the result is
the oracle's point-wise renorm, bit for bit), componentWiseNorm_tree
(wh_component_norm: per-block sums in a fixed tree order, so within rel.
1e-13 of the oracle's sequential sums; Linfty exact), and adapt_tree with
eps_normalized and each eps_norm: grids and fields identical to the
oracle's."""
import numpy as np
import pytest

from helpers import random_graded_forest, smooth_blocks
from oracle import pyoracle as po
from wabbit_amd import lib as L
from wabbit_amd import wavelets as W
from wabbit_amd.tree import Forest

pytestmark = pytest.mark.gpu

KINDS = [W.WH_NORM_LINFTY, W.WH_NORM_L1, W.WH_NORM_L2, W.WH_NORM_H1]


def _pool(dim, bs, g, neq, nb):
    return L.Pool(dim, [bs] * dim, g, neq, nb, n_work_slots=2, n_tmp_eqn=neq)


@pytest.mark.parametrize("dim", [2, 3])
def test_detail_norm_matches_oracle(dim):
    g, bs, neq, n = 6, 16 if dim == 2 else 12, 4, 7
    rng = np.random.default_rng(21)
    nz = bs + 2 * g if dim == 3 else 1
    u = rng.standard_normal((n, neq, nz, bs + 2 * g, bs + 2 * g)) * 10.0 ** rng.integers(-3, 3, (n, neq, 1, 1, 1))
    levels = np.array([0, 1, 2, 3, 4, 5, 7], np.int32)
    pool = _pool(dim, bs, g, neq, n)
    pool.upload(L.WH_BLOCK, 0, u)
    ids = np.array([6, 2, 0, 5, 1, 3, 4], np.int32)
    for kind in KINDS:
        for groups in ([1, 1, 1, 1], [2, 2, 1, 0], [2, 2, 2, 1]):
            print(f"detail dim={dim} eps_norm={kind} groups={groups}", flush=True)
            got = W.detail_blocks_norm(pool, L.WH_BLOCK, ids, levels, groups, neq, kind)
            for k, b in enumerate(ids):
                want = po.detail_block_norm(dim, [bs] * dim, g, u[b], groups, int(levels[k]), kind)
                assert np.array_equal(got[k], want), (kind, groups, k)
    pool.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_component_norm_matches_oracle(dim):
    g, bs, neq, n = 4, 16 if dim == 2 else 8, 3, 9
    rng = np.random.default_rng(22)
    nz = bs + 2 * g if dim == 3 else 1
    u = rng.standard_normal((n, neq, nz, bs + 2 * g, bs + 2 * g))
    domain = [1.0, 2.0, 0.5][:dim]
    levels = rng.integers(1, 5, n).astype(np.int32)
    ids = rng.permutation(n).astype(np.int32)
    pool = _pool(dim, bs, g, neq, n)
    pool.upload(L.WH_BLOCK, 0, u)
    dxs = [np.array([2.0 ** -int(l) * domain[d] / bs for d in range(dim)]) for l in levels]
    for kind in KINDS:
        for groups in ([1, 1, 1], [2, 2, 1], [2, 2, 0]):
            print(f"component norm dim={dim} eps_norm={kind} groups={groups}", flush=True)
            got = W.component_norm(pool, L.WH_BLOCK, ids, levels, groups, neq, kind, domain)
            want = po.component_norm(dim, [bs] * dim, g, [u[b] for b in ids], dxs, groups, kind)
            if kind == W.WH_NORM_LINFTY:
                assert np.array_equal(got, want)
            else:
                assert np.allclose(got, want, rtol=1e-13, atol=0), (kind, groups, got, want)
    pool.close()


# eps chosen per case so that the grid differs from the Linfty one (the
# oracle: 2D 43 blocks -> L2 16 / Linfty 19, L1 10 / 16, CDF40 L2 10 / 19,
# H1 19 / 16; 3D 29 blocks -> L2 and L1 8 / 29, H1 8 at 2e-2 and 29 at 1.8e-2)
@pytest.mark.parametrize("wav,dim,bs,eps,kind", [("CDF44", 2, 16, 1e-4, W.WH_NORM_L2),
                                                 ("CDF44", 2, 16, 1e-3, W.WH_NORM_L1),
                                                 ("CDF40", 2, 24, 1e-2, W.WH_NORM_L2),
                                                 ("CDF44", 2, 16, 1e-3, W.WH_NORM_H1),
                                                 ("CDF44", 3, 16, 3e-4, W.WH_NORM_L2),
                                                 ("CDF44", 3, 16, 3e-4, W.WH_NORM_L1),
                                                 ("CDF44", 3, 16, 2e-2, W.WH_NORM_H1)])
def test_adapt_eps_norm_matches_oracle(wav, dim, bs, eps, kind):
    """adapt_tree with the threshold indicator, eps_normalized and eps_norm
    (details renormed per level, norm of the leaves in that norm, security
    zone for the lifted wavelets): block lists and fields equal to the
    oracle's."""
    w = W.setup(wav)
    g = w.g
    jmin, jmax = 1, 4 if dim == 2 else 3
    tcs, lev = random_graded_forest(dim, jmin, jmax, 3, 8 if dim == 2 else 3)
    n = tcs.shape[0]
    neq = dim + 1
    groups = [2] * dim + [1]
    hvy = smooth_blocks(tcs, lev, jmax, dim, bs, g, 4, neq)
    print(f"adapt {wav} dim={dim} Bs={bs} eps={eps} eps_norm={kind}: {n} blocks", flush=True)
    t = po.Tree(po.wavelet(wav), dim, jmin, jmax, [bs] * dim, g, tcs, lev, hvy, indicator=1, eps=eps,
                groups=groups, eps_normalized=True, eps_norm=kind)
    t.sync()
    t.adapt()
    tco, lvo, hvo = t.get()
    nb_max = 2 * n + 64
    f = Forest(dim, jmax, [bs] * dim, g, [1.0] * dim, nb_max=nb_max)
    f.set_blocks(tcs, lev, np.arange(n, dtype=np.int32))
    pool = _pool(dim, bs, g, neq, nb_max)
    W.set_wavelet(pool, w)
    pool.upload(L.WH_BLOCK, 0, hvy)
    W.sync_tree(pool, f)
    W.adapt_tree(pool, f, W.adapt_params(jmin, jmax, W.WH_IND_THRESHOLD, use_ce=bool(w.lifted), eps=eps,
                                         groups=groups, neq=neq, eps_normalized=True, eps_norm=kind))
    tc, lv, hv, _ = f.blocks()
    key = {(int(a), int(b)): int(h) for a, b, h in zip(tc, lv, hv)}
    print(f"  {n} -> {len(tco)} blocks (device {len(key)})", flush=True)
    assert sorted(key) == sorted((int(a), int(b)) for a, b in zip(tco, lvo))
    assert len(tco) < n
    got = np.stack([pool.download(L.WH_BLOCK, key[(int(a), int(b))], 1)[0] for a, b in zip(tco, lvo)])
    assert np.array_equal(got, hvo)
    pool.close()
    f.close()
