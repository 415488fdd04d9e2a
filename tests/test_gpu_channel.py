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

NSTEPS = 12


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
    states = []
    C.oracle_run(case, tc_o, lv_o, hvy_o, NSTEPS, lambda tree, t, it: states.append((t,) + tree.get()))
    s = wl.channel_solver(case, path=PATH_FUSED, nb_max=8192)
    try:
        for k in range(NSTEPS):
            s.step()
            t, tc, lv, hvy = states[k]
            assert abs(s.time - t) <= 1e-12 * t
            _compare(s, tc, lv, hvy, 1e-12 if k == 0 else 1e-9)
    finally:
        s.close()
