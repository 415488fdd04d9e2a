"""GPU parity: the HIP kernels (through the C ABI) against the oracle and
against the reference's own regression outputs.

Bars: relative L2 <= 1e-12 after one RHS evaluation and one RK step,
<= 1e-9 after the configured step count; block dt,
periodic ghost fill and the step count to t_end must agree exactly.  The
kernels are built with -ffp-contract=off and follow the reference's term
order, so the tests also record whether results are bit-identical.
"""
import ctypes

import numpy as np
import pytest

from helpers import (interior, oracle_run, random_mask, rel_l2, reorder_to_forest, smooth_field,
                     tg3d_case, vortex2d_case)
from wabbit_amd import lib as L
from wabbit_amd.forest import EquidistantForest
from wabbit_amd.params import ACMParams, TimeParams
from wabbit_amd.timestep import PATH_FUSED, PATH_PER_BLOCK, PATH_REFERENCE, AcmSolver

pytestmark = pytest.mark.gpu

SCHEMES = [2, 4, 40, 6]
REACH = {2: 1, 4: 2, 40: 3, 6: 3}


def _oracle():
    from oracle import pyoracle as po
    return po


def _setup(dim, scheme, skew, penal=False, sponge=False, J=None, bs=None):
    J = J if J is not None else (2 if dim == 2 else 1)
    bs = bs if bs is not None else (16 if dim == 2 else 12)
    g = REACH[scheme] + 1
    forest = EquidistantForest(dim, J, [bs] * dim, g, [1.0, 1.5, 2.0][:dim])
    acm = ACMParams(dim=dim, scheme=scheme, skew_symmetry=skew, penalization=penal,
                    use_sponge=sponge, c_0=20.0, nu=1e-3, gamma_p=1.0, C_eta=1e-3,
                    C_sponge=1e-2, u_mean_set=[1.0, 0.2, -0.1], CFL=0.5,
                    domain_size=[1.0, 1.5, 2.0])
    hvy = smooth_field(forest, acm.neq, seed=7 + dim + scheme)
    mask = random_mask(forest, 3) if (penal or sponge) else None
    return forest, acm, hvy, mask


def _oracle_sync(forest, acm, hvy, gs):
    po = _oracle()
    h = hvy.copy()
    bs = np.array(forest.bs, np.int32)
    bat = forest.block_at()
    po.lib().or_sync_equidistant(forest.dim, bs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                 forest.g, acm.neq, forest.nb,
                                 bat.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), forest.n,
                                 h.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), gs)
    return h


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("skew", [False, True])
def test_rhs_one_evaluation(dim, scheme, skew):
    po = _oracle()
    penal = sponge = (scheme == 4)
    forest, acm, hvy, mask = _setup(dim, scheme, skew, penal, sponge)
    g_rhs = REACH[scheme]
    hs = _oracle_sync(forest, acm, hvy, g_rhs)
    pool = L.Pool(dim, forest.bs, forest.g, acm.neq, forest.n, n_work_slots=2,
                  n_mask=6 if mask is not None else 0)
    pool.set_active(forest.active, forest.x0dx)
    pool.upload(L.WH_BLOCK, 0, hs)
    if mask is not None:
        pool.upload(L.WH_MASK, 0, mask)
    pool.rhs(L.acm_params_c(acm), 0.0, L.WH_BLOCK, 1)
    got = interior(pool.download(L.WH_WORK + 1, 0, forest.n), forest.bs[0], forest.g, dim)
    p = po.make_params(acm)
    want = np.stack([
        po.rhs_block(p, forest.bs[:dim], forest.g, forest.x0dx[k, 3:3 + dim], hs[k],
                     None if mask is None else mask[k]) for k in range(forest.n)])
    want = interior(want, forest.bs[0], forest.g, dim)
    for c in range(acm.neq):
        assert rel_l2(got[:, c], want[:, c]) <= 1e-12, (c, rel_l2(got[:, c], want[:, c]))
    print("bit-identical:", np.array_equal(got, want))
    pool.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_sync_same_level_exact(dim):
    forest, acm, hvy, _ = _setup(dim, 4, False)
    for case, gs in ((0, 2), (1, forest.g)):
        want = _oracle_sync(forest, acm, hvy, gs)
        pool = L.Pool(dim, forest.bs, forest.g, acm.neq, forest.n, n_work_slots=0)
        pool.set_active(forest.active, forest.x0dx)
        pool.set_patches(case, forest.same_level_patches(gs))
        pool.upload(L.WH_BLOCK, 0, hvy)
        pool.sync(L.WH_BLOCK, case)
        got = pool.download(L.WH_BLOCK, 0, forest.n)
        assert np.array_equal(got, want)
        pool.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_dt_reduce_exact(dim):
    po = _oracle()
    forest, acm, hvy, _ = _setup(dim, 4, False)
    pool = L.Pool(dim, forest.bs, forest.g, acm.neq, forest.n, n_work_slots=0)
    pool.set_active(forest.active, forest.x0dx)
    pool.upload(L.WH_BLOCK, 0, hvy)
    got = pool.dt(L.acm_params_c(acm))
    p = po.make_params(acm)
    want = min(po.dt_block(p, forest.bs[:dim], forest.g, forest.x0dx[k, 3:3 + dim], hvy[k])
               for k in range(forest.n))
    assert got == want
    pool.close()


@pytest.mark.parametrize("path", [PATH_REFERENCE, PATH_PER_BLOCK, PATH_FUSED])
@pytest.mark.parametrize("dim", [2, 3])
def test_rk_step_paths(dim, path):
    """One RK4 step through each device path against the oracle step."""
    case = vortex2d_case(4) if dim == 2 else tg3d_case(4)
    forest = EquidistantForest(dim, case["J"], [case["Bs"]] * dim, case["g"], [case["L"]] * dim)
    hvy, _ = reorder_to_forest(case, forest)
    case = dict(case, hvy=hvy, origin=None)
    solver = AcmSolver(forest, case["acm"], case["tp"], case["g_rhs"], path=path)
    solver.load(hvy, case["time0"], case["it0"])
    for _ in range(3):
        solver.step()
    got = interior(solver.state(), case["Bs"], case["g"], dim)
    # oracle on the same (forest-ordered) blocks
    po = _oracle()
    run = po.EquidistantRun(po.make_params(case["acm"]), [case["Bs"]] * dim, case["g"],
                            case["g_rhs"], dim + 1, case["J"], [case["L"]] * dim, hvy.copy(),
                            forest.block_at())
    from wabbit_amd.params import clamp_dt
    t = case["time0"]
    for _ in range(3):
        dt = clamp_dt(run.dt(), t, case["tp"])
        run.step(dt)
        t += dt
    want = interior(run.hvy, case["Bs"], case["g"], dim)
    assert solver.time == t
    assert rel_l2(got, want) <= 1e-12
    print("bit-identical:", np.array_equal(got, want))
    solver.pool.close()


def _golden_gpu(case, use_device_loop):
    dim = case["dim"]
    forest = EquidistantForest(dim, case["J"], [case["Bs"]] * dim, case["g"], [case["L"]] * dim)
    assert sorted(forest.treecode.tolist()) == sorted(case["treecode"].tolist())
    hvy, _ = reorder_to_forest(case, forest)
    solver = AcmSolver(forest, case["acm"], case["tp"], case["g_rhs"], path=PATH_FUSED)
    solver.load(hvy, case["time0"], case["it0"])
    if use_device_loop:
        solver.run_device(case["it1"] - case["it0"] + 2)
    else:
        solver.run()
    assert solver.iteration == case["it1"]
    assert abs(solver.time - case["time1"]) <= 1e-12 * case["time1"]
    got = interior(solver.state(), case["Bs"], case["g"], dim)
    bat = reorder_to_forest(case, forest)[1]
    want = case["golden"][bat]
    err = max(rel_l2(got[:, c], want[:, c]) for c in range(dim + 1))
    solver.pool.close()
    return err, np.array_equal(got, want)


@pytest.mark.parametrize("order", [2, 4, 6])
def test_vortex2d_golden_gpu(order):
    err, same = _golden_gpu(vortex2d_case(order), use_device_loop=True)
    print("rel L2", err, "bit-identical", same)
    assert err <= 1e-9


@pytest.mark.parametrize("order", [2, 4, 6])
def test_tg3d_golden_gpu(order):
    err, same = _golden_gpu(tg3d_case(order), use_device_loop=(order != 4))
    print("rel L2", err, "bit-identical", same)
    assert err <= 1e-9


def test_device_loop_matches_host_loop():
    case = vortex2d_case(4)
    forest = EquidistantForest(2, case["J"], [case["Bs"]] * 2, case["g"], [case["L"]] * 2)
    hvy, _ = reorder_to_forest(case, forest)
    res = []
    for dev in (False, True):
        s = AcmSolver(forest, case["acm"], case["tp"], case["g_rhs"], path=PATH_FUSED)
        s.load(hvy, case["time0"], case["it0"])
        if dev:
            s.run_device(40)
        else:
            s.run(40)
        res.append((s.state(), s.time, s.iteration))
        s.pool.close()
    assert res[0][2] == res[1][2] and res[0][1] == res[1][1]
    assert np.array_equal(res[0][0], res[1][0])


def test_divergence_guard_and_errors():
    forest, acm, hvy, _ = _setup(2, 4, False)
    pool = L.Pool(2, forest.bs, forest.g, acm.neq, forest.n, n_work_slots=3)
    pool.set_active(forest.active, forest.x0dx)
    pool.upload(L.WH_BLOCK, 0, hvy)
    m = pool.divergence_guard()
    assert m == pytest.approx(np.abs(interior(hvy, 16, forest.g, 2)).max(), rel=0, abs=0)
    bad = hvy.copy()
    bad[3, 1, 0, forest.g + 2, forest.g + 5] = 2e12
    pool.upload(L.WH_BLOCK, 0, bad)
    with pytest.raises(L.WabbitHipError) as ei:
        pool.divergence_guard()
    assert ei.value.code == -4
    with pytest.raises(L.WabbitHipError):
        pool.set_active(np.array([0, 0], np.int32), np.tile(forest.x0dx[:1], (2, 1)))
    with pytest.raises(L.WabbitHipError):
        pool.upload(L.WH_BLOCK, forest.n - 1, hvy[:2])
    p = forest.same_level_patches(2)
    p[0]["recv_lo"][0] = 1000
    with pytest.raises(L.WabbitHipError):
        pool.set_patches(0, p)
    pool.close()


@pytest.mark.parametrize("scheme", [4, 6])
@pytest.mark.parametrize("ysplit,zsplit,waves,xcd", [(0, 1, 1, 0), (0, 1, 1, 1), (0, 1, 4, 1), (1, 1, 1, 0), (2, 1, 1, 0),
                                                     (3, 1, 1, 0), (2, 2, 1, 0), (3, 4, 1, 0), (2, 3, 3, 0),
                                                     (3, 1, 3, 0)])
def test_3d_kernel_variants_identical(scheme, ysplit, zsplit, waves, xcd):
    """Every tuning variant of the 3D stage kernel (y-split, z-split, waves
    per SIMD, XCD swizzle of the flat kernel) gives the oracle's RHS bit for
    bit (the autotuner may pick any of them)."""
    po = _oracle()
    forest, acm, hvy, mask = _setup(3, scheme, True, True, True, J=1, bs=24)
    hs = _oracle_sync(forest, acm, hvy, REACH[scheme])
    pool = L.Pool(3, forest.bs, forest.g, acm.neq, forest.n, n_work_slots=2, n_mask=6)
    pool.set_active(forest.active, forest.x0dx)
    pool.upload(L.WH_BLOCK, 0, hs)
    pool.upload(L.WH_MASK, 0, mask)
    pool.set_tuning(L.WH_TUNE_K3D_YSPLIT, ysplit)
    pool.set_tuning(L.WH_TUNE_K3D_ZSPLIT, zsplit)
    pool.set_tuning(L.WH_TUNE_K3D_WAVES, waves)
    pool.set_tuning(L.WH_TUNE_K3D_XCD, xcd)
    assert pool.get_tuning(L.WH_TUNE_K3D_ZSPLIT) == zsplit and pool.get_tuning(L.WH_TUNE_K3D_WAVES) == waves
    pool.rhs(L.acm_params_c(acm), 0.0, L.WH_BLOCK, 1)
    got = interior(pool.download(L.WH_WORK + 1, 0, forest.n), 24, forest.g, 3)
    p = po.make_params(acm)
    want = interior(np.stack([po.rhs_block(p, [24] * 3, forest.g, forest.x0dx[k, 3:], hs[k], mask[k])
                              for k in range(forest.n)]), 24, forest.g, 3)
    assert rel_l2(got, want) <= 1e-12
    assert np.array_equal(got, want)
    best, ms = pool.autotune(L.acm_params_c(acm), reps=2)
    assert best in (0, 1, 2, 3, 4, 6, 8) and ms > 0
    assert pool.get_tuning(L.WH_TUNE_K3D_ZSPLIT) in (1, 2, 3, 4, 6)
    pool.close()


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("path", [PATH_FUSED, PATH_REFERENCE])
def test_2d_kernel_variants_identical(scheme, path):
    """Every 2D stage-kernel variant (WH_TUNE_K2D_STAGE 0..4: plain, early
    loads, persistent pipelined with 2/4/8 workgroups per CU) gives the same
    state bit for bit, penalised + sponge, one RK4 step."""
    forest, acm, hvy, mask = _setup(2, scheme, False, True, True)
    out = {}
    for stage in (0, 1, 2, 3, 4):
        solver = AcmSolver(forest, acm, TimeParams(time_max=10.0), REACH[scheme], path=path, mask=mask,
                           n_mask=6)
        L.check(L.load().wh_set_tuning(solver.pool.h, L.WH_TUNE_K2D_STAGE, stage))
        solver.load(hvy)
        solver.step()
        out[stage] = solver.state()
        solver.pool.close()
    for stage in (1, 2, 3, 4):
        assert np.array_equal(out[0], out[stage]), stage
