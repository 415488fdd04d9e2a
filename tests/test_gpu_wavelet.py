"""GPU: CDF wavelet kernels and adapt_tree through the C ABI, against the
oracle (bit-identical: same chains of products, -ffp-contract=off) and the
reference's --coarsen-everywhere outputs (block lists bit-exact, fields
within rel. L2 1e-14 -- those files carry ~1 ulp rounding noise of their
own; see tests/test_oracle_wavelet.py)."""
import numpy as np
import pytest

from helpers import fixture_to_hvy, load_npz, random_graded_forest, smooth_blocks
from oracle import pyoracle as po
from wabbit_amd import lib as L
from wabbit_amd import wavelets as W
from wabbit_amd.tree import Forest, refine_on_device

pytestmark = pytest.mark.gpu

WAVS = ["CDF20", "CDF22", "CDF40", "CDF42", "CDF44", "CDF60", "CDF62"]


def _pool(dim, bs, g, neq, nb):
    return L.Pool(dim, [bs] * dim, g, neq, nb, n_work_slots=2, n_tmp_eqn=neq)


def _block_transforms(wav, dim, bs, n=5):
    w = W.setup(wav)
    wo = po.wavelet(wav)
    g = w.g
    neq = 2
    rng = np.random.default_rng(11)
    nz = bs + 2 * g if dim == 3 else 1
    shape = (n, neq, nz, bs + 2 * g, bs + 2 * g)
    src = rng.standard_normal(shape)
    prior = rng.standard_normal(shape)
    pool = _pool(dim, bs, g, neq, n + 2)
    W.set_wavelet(pool, w)
    ids = np.array([3, 0, 4, 1, 2] if n == 5 else list(range(n - 1, -1, -1)), np.int32)   # a permuted list
    for op, ofn in (("fwt", po.fwt_block), ("iwt", po.iwt_block), ("filter", None)):
        pool.upload(L.WH_TMP, 0, src)
        pool.upload(L.WH_BLOCK, 0, prior)
        getattr(W, f"{op}_blocks")(pool, L.WH_TMP, L.WH_BLOCK, ids)
        got = pool.download(L.WH_BLOCK, 0, n)
        for b in range(n):
            if op == "filter":
                want = po.filter_block(wo, dim, [bs] * dim, g, src[b])
            else:
                want = ofn(wo, dim, [bs] * dim, g, src[b], prior[b])
            assert np.array_equal(got[b], want), (op, b)
    pool.close()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("wav", WAVS + ["CDF64", "CDF66"])
def test_block_transforms_match_oracle(wav, dim):
    """FWT, IWT and the HD filter of a permuted block list, bit for bit.
    CDF64/CDF66 are the wavelets whose reconstruction is the general chain
    with breaks (two sums added), and whose decomposition chains are longer
    than the unrolled kernel variants."""
    _block_transforms(wav, dim, 16 if dim == 2 else 12)


@pytest.mark.parametrize("wav,dim,bs", [("CDF20", 2, 64), ("CDF44", 2, 64), ("CDF66", 2, 60),
                                        ("CDF44", 3, 54)])
def test_block_transforms_large_blocks(wav, dim, bs):
    """Blocks whose planes do not fit the fused x/y kernel's 64 KB of LDS
    take the pass-per-launch path (x pass per output point, y and z passes
    in place through the column ring): bit for bit as well."""
    _block_transforms(wav, dim, bs, n=3 if dim == 3 else 5)


@pytest.mark.parametrize("dim", [2, 3])
def test_detail_matches_oracle(dim):
    g, bs, neq, n = 6, 16 if dim == 2 else 12, 4, 6
    rng = np.random.default_rng(5)
    nz = bs + 2 * g if dim == 3 else 1
    u = rng.standard_normal((n, neq, nz, bs + 2 * g, bs + 2 * g))
    pool = _pool(dim, bs, g, neq, n)
    pool.upload(L.WH_BLOCK, 0, u)
    for groups in ([1, 1, 1, 1], [2, 2, 1, 0], [2, 2, 2, 1]):
        got = W.detail_blocks(pool, L.WH_BLOCK, np.arange(n), groups, neq)
        for b in range(n):
            want = po.detail_block(dim, [bs] * dim, g, u[b], groups)
            assert np.array_equal(got[b], want)
    pool.close()


def test_boxes():
    dim, g, bs, neq, n = 2, 4, 16, 2, 3
    rng = np.random.default_rng(3)
    u = rng.standard_normal((n, neq, 1, bs + 2 * g, bs + 2 * g))
    t = rng.standard_normal(u.shape)
    pool = _pool(dim, bs, g, neq, n)
    pool.upload(L.WH_BLOCK, 0, u)
    pool.upload(L.WH_TMP, 0, t)
    boxes = np.zeros(3, W.BOX_DTYPE)
    boxes[0] = (1, (0, 2, 0), (5, 9, 0), W.WH_BOX_ZERO_WC)
    boxes[1] = (2, (g, g, 0), (g + 7, g + bs - 1, 0), W.WH_BOX_COPY_SC)
    boxes[2] = (1, (g + bs, 0, 0), (bs + 2 * g - 1, bs + 2 * g - 1, 0), W.WH_BOX_ZERO_WC)
    W.apply_boxes(pool, L.WH_BLOCK, L.WH_TMP, boxes)
    got = pool.download(L.WH_BLOCK, 0, n)
    want = u.copy()
    for b in boxes:
        sl = (b["hvy"], slice(None), slice(b["lo"][2], b["hi"][2] + 1), slice(b["lo"][1], b["hi"][1] + 1),
              slice(b["lo"][0], b["hi"][0] + 1))
        iy = np.arange(b["lo"][1], b["hi"][1] + 1)[:, None]
        ix = np.arange(b["lo"][0], b["hi"][0] + 1)[None, :]
        sc = (((ix - g) % 2) == 0) & (((iy - g) % 2) == 0)
        blk = want[sl]
        if b["op"] == W.WH_BOX_ZERO_WC:
            blk[..., ~sc] = 0.0
        else:
            blk[..., sc] = t[sl][..., sc]
        want[sl] = blk
    assert np.array_equal(got, want)
    pool.close()


def _graded_case(dim, jmin, jmax, bs, g, seed, nref, neq=1):
    tcs, lev = random_graded_forest(dim, jmin, jmax, seed, nref)
    n = tcs.shape[0]
    rng = np.random.default_rng(seed)
    nz = bs + 2 * g if dim == 3 else 1
    hvy = rng.standard_normal((n, neq, nz, bs + 2 * g, bs + 2 * g))
    return tcs, lev, hvy


@pytest.mark.parametrize("wav,dim,bs", [("CDF44", 2, 16), ("CDF22", 2, 16), ("CDF62", 2, 24),
                                        ("CDF40", 2, 16), ("CDF44", 3, 12)])
def test_sync_tree_matches_oracle(wav, dim, bs):
    w = W.setup(wav)
    g = w.g
    jmax = 4 if dim == 2 else 3       # the treecodes' encoding level
    tcs, lev, hvy = _graded_case(dim, 1, jmax, bs, g, 7, 6 if dim == 2 else 3, neq=2)
    n = tcs.shape[0]
    f = Forest(dim, jmax, [bs] * dim, g, [1.0] * dim, nb_max=n)
    f.set_blocks(tcs, lev, np.arange(n, dtype=np.int32))
    pool = _pool(dim, bs, g, 2, n)
    W.set_wavelet(pool, w)
    pool.upload(L.WH_BLOCK, 0, hvy)
    W.sync_tree(pool, f)
    got = pool.download(L.WH_BLOCK, 0, n)
    t = po.Tree(po.wavelet(wav), dim, 1, jmax, [bs] * dim, g, tcs, lev, hvy)
    t.sync()
    tc2, lv2, want = t.get()
    pos = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(tcs, lev))}
    assert sorted(pos) == sorted((int(a), int(b)) for a, b in zip(tc2, lv2))
    got = np.stack([got[pos[(int(a), int(b))]] for a, b in zip(tc2, lv2)])
    assert np.array_equal(got, want)
    pool.close()
    f.close()


def _gpu_coarsen_everywhere(wav):
    w = W.setup(wav)
    g = w.g
    src = load_npz("wavelet_input.npz")
    bs, jmax = int(src["bs"][0]), int(src["max_level"])
    n = src["treecode"].shape[0]
    hvy = fixture_to_hvy([src["field_vor"]], src["origin"], src["spacing"], bs, g, 2)
    nb_max = 4 * n + 300
    f = Forest(2, jmax, [bs, bs], g, [float(src["domain_size"][0])] * 2, nb_max=nb_max)
    f.set_blocks(src["treecode"], src["level"], np.arange(n, dtype=np.int32))
    pool = _pool(2, bs, g, 1, nb_max)
    W.set_wavelet(pool, w)
    f.attach(pool, g, w.order)
    pool.upload(L.WH_BLOCK, 0, hvy)
    # wabbit-post --refine-everywhere (read, sync_ghosts_tree, refine) ...
    refine_on_device(pool, f, np.arange(n, dtype=np.int32), w.order, g, presync="tree")
    # ... then --coarsen-everywhere (read, sync_ghosts_tree, adapt_tree)
    W.sync_tree(pool, f)
    W.adapt_tree(pool, f, W.adapt_params(1, jmax, W.WH_IND_EVERYWHERE, use_ce=bool(w.lifted)))
    tc, lv, hv, _ = f.blocks()
    data = np.stack([pool.download(L.WH_BLOCK, int(h), 1)[0] for h in hv])
    pool.close()
    f.close()
    return g, bs, tc, lv, data


@pytest.mark.parametrize("wav", WAVS)
def test_coarsen_everywhere_gpu(wav):
    g, bs, tc, lv, data = _gpu_coarsen_everywhere(wav)
    # the oracle, bit for bit
    from test_oracle_wavelet import coarsen_everywhere
    g2, bs2, tco, lvo, hvo = coarsen_everywhere(wav)
    mine = {(int(t), int(l)): k for k, (t, l) in enumerate(zip(tc, lv))}
    assert sorted(mine) == sorted((int(t), int(l)) for t, l in zip(tco, lvo))
    got = np.stack([data[mine[(int(t), int(l))]] for t, l in zip(tco, lvo)])
    assert np.array_equal(got, hvo)
    # the reference's files
    src = load_npz("wavelet_input.npz")
    ref = src if wav in ("CDF20", "CDF40", "CDF60") else load_npz(f"wavelet_coarsen_{wav}.npz")
    out = np.stack([data[mine[(int(t), int(l))]][0, 0, g:g + bs + 1, g:g + bs + 1]
                    for t, l in zip(ref["treecode"], ref["level"])])
    want = ref["field_vor"]
    if wav in ("CDF20", "CDF40", "CDF60"):
        assert np.array_equal(out[:, :bs, :bs], want[:, :bs, :bs])
    else:
        assert np.linalg.norm(out - want) / np.linalg.norm(want) <= 1e-14


@pytest.mark.parametrize("wav,dim,bs,eps", [("CDF44", 2, 16, 1e-3), ("CDF44", 2, 24, 1e-3),
                                            ("CDF22", 2, 16, 1e-2), ("CDF40", 2, 16, 1e-3),
                                            ("CDF62", 2, 24, 1e-3), ("CDF44", 3, 16, 1e-2)])
def test_adapt_threshold_matches_oracle(wav, dim, bs, eps):
    """adapt_tree with the threshold indicator on a random graded forest:
    flags, block lists and fields identical to the oracle (covers leaf-first
    and level-wise decomposition, leaf and level-wise reconstruction)."""
    w = W.setup(wav)
    g = w.g
    jmin, jmax = 1, 4 if dim == 2 else 3
    tcs, lev = random_graded_forest(dim, jmin, jmax, 3, 8 if dim == 2 else 3)
    n = tcs.shape[0]
    neq = 2
    hvy = smooth_blocks(tcs, lev, jmax, dim, bs, g, 4, neq)
    t = po.Tree(po.wavelet(wav), dim, jmin, jmax, [bs] * dim, g, tcs, lev, hvy, indicator=1, eps=eps,
                groups=[1, 1])
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
                                         groups=[1, 1]))
    tc, lv, hv, _ = f.blocks()
    key = {(int(a), int(b)): int(h) for a, b, h in zip(tc, lv, hv)}
    assert sorted(key) == sorted((int(a), int(b)) for a, b in zip(tco, lvo))
    assert len(tco) < n or wav == "CDF40"           # something was coarsened
    got = np.stack([pool.download(L.WH_BLOCK, key[(int(a), int(b))], 1)[0] for a, b in zip(tco, lvo)])
    assert np.array_equal(got, hvo)
    pool.close()
    f.close()
