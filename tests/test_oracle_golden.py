"""The oracle against the reference's own regression outputs.

TESTING/acm/3vortices/3vorticesEquiFD{2,4,6}_CDF{20,40,60} (2D, 64 blocks,
Bs=32, skew-symmetric ACM, RK4 from t=10 to t=20: 3073 steps) and
TESTING/acm/taylorGreen/taylorGreenEqui_FD{2,4,6}_CDF{20,40,60} (3D, 8 blocks,
Bs=20, t=0..10 with dt truncated at the 0.2 statistics interval).  These pin
the RHS (each scheme, skew form), the block dt, RK4, the periodic same-level
ghost fill and the host dt clamps.  The reference was built with gfortran
-O3 (no FMA on x86-64); the oracle reproduces its outputs bit for bit.
"""
import numpy as np
import pytest

from helpers import interior, oracle_run, rel_l2, tg3d_case, vortex2d_case


@pytest.mark.slow
@pytest.mark.parametrize("order", [2, 4, 6])
def test_vortex2d_golden(order):
    case = vortex2d_case(order)
    hvy, t, it = oracle_run(case)
    assert it == case["it1"] and t == case["time1"]
    mine = interior(hvy, case["Bs"], case["g"], 2)
    for c in range(3):
        assert rel_l2(mine[:, c], case["golden"][:, c]) <= 1e-12
    assert np.array_equal(mine, case["golden"])  # bit-identical


@pytest.mark.slow
@pytest.mark.parametrize("order", [2, 4, 6])
def test_tg3d_golden(order):
    case = tg3d_case(order)
    hvy, t, it = oracle_run(case)
    assert it == case["it1"] and abs(t - case["time1"]) == 0.0
    mine = interior(hvy, case["Bs"], case["g"], 3)
    for c in range(4):
        assert rel_l2(mine[:, c], case["golden"][:, c]) <= 1e-12
    assert np.array_equal(mine, case["golden"])


def test_planted_mistake_is_caught():
    """A plausible oracle mistake (wrong sign of nu) must fail the golden check
    within a few steps: run 20 steps with nu flipped and compare against the
    correct oracle run."""
    case = vortex2d_case(4)
    ref, _, _ = oracle_run(case, nsteps=20)
    case["acm"].nu = -case["acm"].nu
    bad, _, _ = oracle_run(case, nsteps=20)
    a = interior(ref, case["Bs"], case["g"], 2)
    b = interior(bad, case["Bs"], case["g"], 2)
    assert rel_l2(b, a) > 1e-9
