"""GPU: the batched RK step (wh_batch_rk_step) while its entry count grows
from step to step, so its device tables are re-grown between steps.

The entry, group, dt and origin/spacing tables grow by different factors;
the sequence 1024 -> 1104 -> 1552 blocks grows them once and then needs
more origin/spacing room than the first growth gave (6 doubles per entry)
while the entry table still fits.  Every step must equal the one-pool
grouped step (wh_rk_step_groups, fused path) on an identical pool, bit for
bit, with and without graph replay and for every variant of the batch's 2D
stage kernel.
"""
import numpy as np
import pytest

from wabbit_amd import lib as L
from wabbit_amd.forest import EquidistantForest
from wabbit_amd.params import ACMParams, butcher_rk4

pytestmark = pytest.mark.gpu

BS, G, NB = 16, 3, 2048
# independent periodic forests (levels) packed into one pool, per step
STEPS = [(5,), (5, 3, 2), (5, 4, 4, 2)]


def _tables(levels):
    act, x0dx, grp, pts = [], [], [], []
    base = 0
    for gi, lev in enumerate(levels):
        f = EquidistantForest(2, lev, [BS, BS], G, [1.0, 1.0])
        act.append(f.active + base)
        x0dx.append(f.x0dx)
        grp.append(np.full(f.n, gi, np.int32))
        p = f.same_level_patches(2).copy()
        p["recv"] += base
        p["send"] += base
        pts.append(p)
        base += f.n
    return (np.concatenate(act).astype(np.int32), np.concatenate(x0dx), np.concatenate(grp),
            np.concatenate(pts), len(levels))


@pytest.mark.parametrize("use_graph,variant", [(False, 0), (True, 0), (True, 1), (True, 2), (True, 3), (True, 4)])
def test_batch_step_with_growing_tables(use_graph, variant):
    acm = ACMParams(dim=2, scheme=4, skew_symmetry=False, penalization=False, c_0=20.0, nu=1e-3,
                    gamma_p=1.0, CFL=0.5, domain_size=[1.0, 1.0, 1.0])
    acm_c = L.acm_params_c(acm)
    bt = butcher_rk4()
    rng = np.random.default_rng(7)
    a = L.Pool(2, [BS, BS], G, 3, NB, n_work_slots=3)
    b = L.Pool(2, [BS, BS], G, 3, NB, n_work_slots=3)
    st = L.BatchStepper([a])
    st.set_tuning(L.WH_TUNE_K2D_STAGE, variant)   # every 2D stage-kernel variant: the same bits
    try:
        data = rng.standard_normal((NB,) + a.block_shape()) * 0.1
        a.upload(L.WH_BLOCK, 0, data)
        b.upload(L.WH_BLOCK, 0, data)
        for levels in STEPS:
            act, x0dx, grp, pts, ng = _tables(levels)
            for p in (a, b):
                p.set_active(act, x0dx)
                p.set_patches(0, pts)
                p.set_block_groups(grp, ng)
            dts = 1e-5 * (1.0 + np.arange(ng))
            print(f"batch step over {act.size} blocks in {ng} groups, graph={use_graph}", flush=True)
            st.rk_step(acm_c, bt, dts, use_graph=use_graph)
            st.synchronize()
            b.rk_step_groups(acm_c, bt, dts, path=2)
            got = a.download(L.WH_BLOCK, 0, NB)
            want = b.download(L.WH_BLOCK, 0, NB)
            inner = (act, slice(None), 0, slice(G, G + BS), slice(G, G + BS))
            assert np.array_equal(got[inner], want[inner]), levels
            assert not np.array_equal(got[inner], data[inner]), levels
    finally:
        st.close()
        a.close()
        b.close()
