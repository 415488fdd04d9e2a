"""GPU: wh_fwt_threshold_noalloc (the entry point that allocates nothing)
against wh_fwt_threshold, through the C ABI.

Two calls in a row on the same pool pair, the second with fewer blocks and
new data, in 2D and 3D: flags, details and the decomposed blocks must equal
the allocating entry point's, bit for bit (the same kernels in the same
order; the second call must not see the first call's ids or details).  A
call with more than WH_NOALLOC_MAX_BLOCKS blocks is refused before any
device work."""
import numpy as np
import pytest

from wabbit_amd import lib as L
from wabbit_amd import wavelets as W

pytestmark = pytest.mark.gpu

MAX_BLOCKS = 16384  # WH_NOALLOC_MAX_BLOCKS in include/wabbit_hip.h


def _blocks(rng, n, dim, bs, g, neq):
    nz = bs + 2 * g if dim == 3 else 1
    x = np.arange(bs + 2 * g) / (bs + 2 * g)
    base = np.sin(2 * np.pi * x)[None, :] * np.cos(2 * np.pi * x)[:, None]
    return np.stack([np.stack([(0.01 * (k % 7 + 1)) * base[None] + 1e-3 * (k % 3) * rng.standard_normal((nz,) + base.shape)
                               for _ in range(neq)]) for k in range(n)])


@pytest.mark.parametrize("dim,bs,nb,calls", [(2, 16, 400, (300, 37)), (3, 16, 48, (40, 9))])
def test_noalloc_matches_fwt_threshold(dim, bs, nb, calls):
    w = W.setup("CDF44")
    g, neq = w.g, dim + 1
    rng = np.random.default_rng(11 + dim)
    pools = []
    try:
        for _ in range(2):
            p = L.Pool(dim, [bs] * dim, g, neq, nb, n_work_slots=1, n_tmp_eqn=neq)
            W.set_wavelet(p, w)
            pools.append(p)
        a, b = pools
        groups = [1] * (neq - 1) + [0] if dim == 2 else [2, 2, 2, 1]
        for n in calls:
            src = _blocks(rng, nb, dim, bs, g, neq)
            for p in pools:
                p.upload(L.WH_TMP, 0, src)
                p.upload(L.WH_BLOCK, 0, src)
            ids = rng.permutation(nb)[:n].astype(np.int32)
            print(f"dim {dim}: call with {n} blocks", flush=True)
            f_ref, d_ref = W.fwt_threshold(a, L.WH_TMP, L.WH_BLOCK, ids, np.zeros(neq), groups)
            thr = np.median(d_ref, axis=0)
            f_ref, d_ref = W.fwt_threshold(a, L.WH_TMP, L.WH_BLOCK, ids, thr, groups)
            f_got, d_got = W.fwt_threshold_noalloc(b, L.WH_TMP, L.WH_BLOCK, ids, thr, groups)
            assert np.array_equal(d_got, d_ref)
            assert np.array_equal(f_got, f_ref)
            assert 0 < (f_got == -1).sum() < n
            assert np.array_equal(b.download(L.WH_BLOCK, 0, nb), a.download(L.WH_BLOCK, 0, nb))
        # beyond the fixed array: refused with an error, nothing launched
        with pytest.raises(L.WabbitHipError, match="WH_NOALLOC_MAX_BLOCKS"):
            W.fwt_threshold_noalloc(b, L.WH_TMP, L.WH_BLOCK, np.zeros(MAX_BLOCKS + 1, np.int32), np.zeros(neq),
                                    groups)
    finally:
        for p in pools:
            p.close()
