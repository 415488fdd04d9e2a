"""GPU: the adaptive 3D periodic workload (CDF44, eps=1e-3, solenoidal
Fourier velocity) at reduced size (Bs=16, Jmin=2, Jmax=3) on the device
against the oracle: the initial grid and two main-loop iterations (sync,
refine everywhere, RK4, adapt_tree).

Bars: block lists identical after every adapt_tree; fields within relative
L2 1e-12 after the first step and 1e-9 after the second."""
import numpy as np
import pytest

import periodic3d_case as C
from wabbit_amd import workloads as wl

pytestmark = pytest.mark.gpu

NSTEPS = 2


def test_adaptive_3d_matches_oracle():
    w = wl.periodic3d(jmax=3, bs=16)
    tc_o, lv_o, hvy_o = C.oracle_initial_grid(w)
    print(f"initial grid {len(tc_o)} blocks", flush=True)
    states = []
    C.oracle_run(w, tc_o, lv_o, hvy_o, NSTEPS, lambda tree, t, it: states.append((t,) + tree.get()))
    s = wl.adaptive_solver(w, nb_max=4096)
    try:
        tc, lv, data = s.state()
        assert np.array_equal(tc, tc_o) and np.array_equal(lv, lv_o)
        assert np.array_equal(data, hvy_o)
        g, bs = C.G, w["bs"]
        inner = (slice(None), slice(None), slice(g, g + bs), slice(g, g + bs), slice(g, g + bs))
        for k in range(NSTEPS):
            s.step()
            t, tco, lvo, hvo = states[k]
            print(f"step {k + 1}: {len(tco)} blocks", flush=True)
            assert abs(s.time - t) <= 1e-12 * t
            tc, lv, data = s.state()
            assert np.array_equal(tc, tco) and np.array_equal(lv, lvo)
            rel = np.linalg.norm(data[inner] - hvo[inner]) / np.linalg.norm(hvo[inner])
            assert rel <= (1e-12 if k == 0 else 1e-9), rel
    finally:
        s.close()
