"""GPU: coarsening flags of threshold_block on realistic data, with margins.

The leaves of the channel workload after a few main-loop steps are
decomposed (one FWT level) and thresholded on the device and on the oracle
(the reference's threshold_block: max |WC| per component over the interior,
flag -1 iff every component's detail <= eps).  The flags must agree block by
block; a flip would be reported with its detail-coefficient margin
|detail - eps| / eps, and the test also prints the closest margin seen."""
import numpy as np
import pytest

from oracle import pyoracle as po
from wabbit_amd import lib as L
from wabbit_amd import wavelets as W
from wabbit_amd import workloads as wl

pytestmark = pytest.mark.gpu


def test_flags_and_margins_on_channel_state():
    w = wl.channel(jmax=5)
    s = wl.channel_solver(w, nb_max=4096)
    try:
        for _ in range(6):
            s.step()
        tc, lv, data = s.state()
    finally:
        s.close()
    n, neq = len(tc), data.shape[1]
    g, bs = 6, w["bs"]
    wv = W.setup(w["wavelet"], 2)
    pool = L.Pool(2, [bs, bs], g, neq, 2 * n, n_work_slots=2, n_tmp_eqn=neq)
    try:
        W.set_wavelet(pool, wv)
        pool.upload(L.WH_TMP, 0, data)
        pool.upload(L.WH_BLOCK, 0, data)     # the transform keeps dst outside what it computes
        ids = np.arange(n, dtype=np.int32)
        W.fwt_blocks(pool, L.WH_TMP, L.WH_BLOCK, ids)
        det_gpu = W.detail_blocks(pool, L.WH_BLOCK, ids, [1, 1, 1], neq)
        fwt_gpu = pool.download(L.WH_BLOCK, 0, n)
    finally:
        pool.close()
    wo = po.wavelet(w["wavelet"], fd_max_size=2)
    eps = w["eps"]
    flips, closest = [], np.inf
    for b in range(n):
        u_fwt = po.fwt_block(wo, 2, [bs, bs], g, data[b])
        det = po.detail_block(2, [bs, bs], g, u_fwt, [1, 1, 1])
        f_gpu = bool(np.all(det_gpu[b] <= eps))
        f_ora = bool(np.all(det <= eps))
        margin = float(np.min(np.abs(det - eps)) / eps)
        closest = min(closest, margin)
        if f_gpu != f_ora:
            flips.append((int(tc[b]), int(lv[b]), margin))
        assert np.array_equal(det_gpu[b], det), b
    print(f"{n} leaves, closest detail margin to eps: {closest:.3e} (relative)", flush=True)
    assert not flips, f"flag flips (treecode, level, margin): {flips}"
    assert np.isfinite(closest)
