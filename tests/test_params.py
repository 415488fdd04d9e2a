"""Host logic: INI parsing, dt clamps, treecode encoding."""
import numpy as np
import pytest

from helpers import load_npz
from wabbit_amd.forest import EquidistantForest, ijk_from_treecode, treecode_from_ijk
from wabbit_amd.params import IniFile, TimeParams, butcher_rk4, clamp_dt, params_from_ini

INI = """
[Domain]
dim=2;
domain_size=4.0 2.0;
[Blocks]
number_block_nodes=16;
[Time]
time_max=2.5;
CFL=0.5;
write_method=fixed_time;
write_time=1.0;
butcher_tableau=(/0.0 0.0 0.0
0.5 0.5 0.0
0.0 0.0 1.0/);
[ACM-new]
c_0=20; comment after value
nu=1e-3;
skew_symmetry=0;
u_mean_set=1.0 0.0 0.0;
[VPM]
penalization=1;
C_eta=1e-3;
[Sponge]
use_sponge=1;
C_sponge=;
[Discretization]
order_discretization=FD_4th_central;
"""


def test_ini_parse_and_defaults():
    acm, tp = params_from_ini(IniFile(INI))
    assert acm.dim == 2 and acm.scheme == 4 and acm.c_0 == 20.0 and acm.nu == 1e-3
    assert acm.penalization and acm.use_sponge and acm.C_sponge == 1e-2  # empty -> default
    assert acm.CFL == 0.5 and acm.CFL_eta == 0.99
    assert acm.CFL_nu == pytest.approx(0.95 * 2.79 / (5.333 * 2))
    assert tp.write_method == "fixed_time" and tp.time_max == 2.5
    assert tp.butcher.shape == (3, 3) and tp.butcher[1, 1] == 0.5
    assert np.array_equal(butcher_rk4()[4], [0, 1 / 6, 1 / 3, 1 / 3, 1 / 6])


def test_clamp_dt_hits_output_and_end_times():
    tp = TimeParams(time_max=1.0, write_method="fixed_time", write_time=0.25)
    t, n = 0.0, 0
    hits = []
    while t < tp.time_max:
        dt = clamp_dt(0.07, t, tp)
        t += dt
        n += 1
        hits.append(t)
    assert t == 1.0
    for w in (0.25, 0.5, 0.75):
        assert any(abs(h - w) < 1e-14 for h in hits)
    # dt_fixed wins over the physics dt, dt_max caps it
    assert clamp_dt(0.07, 0.0, TimeParams(dt_fixed=0.01, time_max=1.0)) == 0.01
    assert clamp_dt(0.07, 0.0, TimeParams(dt_max=0.05, time_max=1.0)) == 0.05


def test_treecodes_match_reference_files():
    """The equidistant forest's treecodes are the ones the reference wrote
    for the same grid (block_treecode_num of the 3vortices / Taylor-Green files)."""
    for name, dim, J, bs in (("vortex2d_ic.npz", 2, 3, 32), ("tg3d_fd4.npz", 3, 1, 20)):
        f = load_npz(name)
        L = float(f["domain_size"][0])
        forest = EquidistantForest(dim, J, [bs] * dim, 3, [L] * dim)
        assert sorted(forest.treecode.tolist()) == sorted(f["treecode"].tolist())
        # origins too: same treecode -> same origin
        by_tc = {int(t): k for k, t in enumerate(forest.treecode)}
        for n in range(f["treecode"].shape[0]):
            k = by_tc[int(f["treecode"][n])]
            assert np.array_equal(forest.x0dx[k, :dim], f["origin"][n])
            assert np.array_equal(forest.x0dx[k, 3:3 + dim], f["spacing"][n])


def test_treecode_roundtrip():
    for dim in (2, 3):
        for lev in range(0, 6):
            rng = np.random.default_rng(lev)
            for _ in range(20):
                ijk = [int(x) for x in rng.integers(0, 2 ** lev, size=3)]
                if dim == 2:
                    ijk[2] = 0
                tc = treecode_from_ijk(ijk, lev, 7, dim)
                assert ijk_from_treecode(tc, lev, 7, dim) == ijk
