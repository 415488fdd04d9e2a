"""GPU: the adaptive channel-cylinder workload (2D ACM, Bs=16, Jmin=2,
Jmax=6, CDF44, eps=1e-3, adapt every step) on the
device against the oracle: the initial grid of setInitialCondition_tree and
the main loop, step by step.

Bars: block lists (treecodes, levels) identical after every adapt_tree;
fields within relative L2 1e-12 after the first step and 1e-9 after the
last (the device's cos/pow in the mask may differ from the host libm by an
ulp, so fields are not compared bit for bit)."""
import numpy as np
import pytest

import channel_case as C
from wabbit_amd import workloads as wl
from wabbit_amd.timestep import PATH_FUSED

pytestmark = pytest.mark.gpu

NSTEPS = 500


@pytest.fixture(scope="module")
def case():
    w = wl.channel()
    print(f"channel case {w['name']}: cylinder {w['mask']['x_cntr'][:2]} R={w['mask']['R_cyl']:.4f}", flush=True)
    return w


def _compare(s, tc_o, lv_o, hvy_o, tol):
    tc, lv, data = s.state()
    assert np.array_equal(tc, tc_o) and np.array_equal(lv, lv_o)
    g, bs = C.G, s.cfg.bs
    inner = (slice(None), slice(None), 0, slice(g, g + bs), slice(g, g + bs))
    assert C.rel_l2(data[inner], hvy_o[inner]) <= tol


def test_initial_grid(case):
    tc_o, lv_o, hvy_o = C.oracle_initial_grid(case)
    s = wl.channel_solver(case, path=PATH_FUSED, nb_max=8192)
    try:
        _compare(s, tc_o, lv_o, hvy_o, 0.0)
        # the mask (cylinder + rect sponge on the 4 x 2 domain) on this grid
        tree = C.oracle_tree(case, tc_o, lv_o, hvy_o)
        want = tree.get_mask(C.po.mask_params(**case["mask"]))
        got = s.mask()
        assert got.shape == want.shape
        assert np.max(np.abs(got - want)) <= 4e-16      # ulp of the device cos/sqrt at most
    finally:
        s.close()


def test_steps(case):
    tc_o, lv_o, hvy_o = C.oracle_initial_grid(case)
    oracle = C.oracle_steps(case, tc_o, lv_o, hvy_o)
    s = wl.channel_solver(case, path=PATH_FUSED, nb_max=8192)
    try:
        for k in range(NSTEPS):
            s.step()
            t, tc, lv, hvy = next(oracle)
            if (k + 1) % 20 == 0:
                print(f"step {k + 1}: {len(tc)} blocks, t={t:.6f}", flush=True)
            assert abs(s.time - t) <= 1e-12 * t
            _compare(s, tc, lv, hvy, 1e-12 if k == 0 else 1e-9)
    finally:
        oracle.close()
        s.close()


def test_device_time_matches_host_dt(case):
    """The fused step with dt, its clamps and the time advance on the device
    (wh_run_steps + wh_time_fetch/wh_time_fetched) against the same solver
    taking dt on the host (wh_dt_reduce + clamp_dt + wh_rk_step): identical
    dt, time, iteration, grids and fields, bit for bit."""
    out = {}
    for dev in (True, False):
        s = wl.channel_solver(case, path=PATH_FUSED, nb_max=8192)
        s.device_time = dev
        try:
            dts = [s.step() for _ in range(5)]
            out[dev] = (dts, s.time, s.iteration, s.state())
        finally:
            s.close()
    (d1, t1, i1, (tc1, lv1, u1)), (d2, t2, i2, (tc2, lv2, u2)) = out[True], out[False]
    assert d1 == d2 and t1 == t2 and i1 == i2 == 5
    assert np.array_equal(tc1, tc2) and np.array_equal(lv1, lv2) and np.array_equal(u1, u2)


def test_stage_kernel_timing_leaves_state(case):
    """wh_time_stage_kernel times RK4 stage sequences on the live grid and
    writes only work slots: the state and the next step are unchanged."""
    from wabbit_amd import lib as L
    ref = wl.channel_solver(case, path=PATH_FUSED, nb_max=8192)
    s = wl.channel_solver(case, path=PATH_FUSED, nb_max=8192)
    try:
        for x in (ref, s):
            x.step()
        before = s.pool.download(L.WH_BLOCK, 0, s.nb_max)
        ms = s.pool.time_stage_kernel(s.acm_c, 3)
        assert ms > 0.0
        assert np.array_equal(before, s.pool.download(L.WH_BLOCK, 0, s.nb_max))
        for x in (ref, s):
            x.step()
        a, b = ref.state(), s.state()
        assert all(np.array_equal(p, q) for p, q in zip(a, b))
    finally:
        ref.close()
        s.close()
