"""GPU: ghost sync with level jumps, block refinement and RK4 on non-uniform
forests, through the C ABI, against the oracle and the reference's
--refine-everywhere outputs.  Bars: bit-identical where the oracle is
(integer tables, copies, interpolation in the reference's order); relative
L2 <= 1e-12 after one RK4 step."""
import numpy as np
import pytest

from helpers import (fixture_to_hvy, interior, load_npz, params_of, random_block_data,
                     random_graded_forest, rel_l2)
from oracle import pyoracle as po
from wabbit_amd import lib as L
from wabbit_amd.params import ACMParams, TimeParams, butcher_rk4, clamp_dt
from wabbit_amd.tree import Forest, refine_on_device

pytestmark = pytest.mark.gpu

CASES = [(2, 2, 5, 8, 3, 4, 3), (2, 2, 5, 8, 3, 4, 2), (2, 1, 4, 8, 1, 2, 1),
         (2, 2, 5, 12, 5, 6, 5), (3, 1, 3, 6, 3, 4, 3), (3, 1, 3, 8, 3, 4, 2)]


@pytest.mark.parametrize("case", CASES)
def test_sync_level_jumps_exact(case):
    dim, jmin, jmax, bs, g, order, depth = case
    tcs, lev = random_graded_forest(dim, jmin, jmax, 3, nref=6 if dim == 2 else 3)
    n = tcs.shape[0]
    neq = dim + 1
    hvy = random_block_data(n, neq, dim, bs, g, 5)
    f = Forest(dim, jmax, [bs] * dim, g, [1.0, 2.0, 1.5][:dim], nb_max=n)
    f.set_blocks(tcs, lev, np.arange(n, dtype=np.int32))
    pool = L.Pool(dim, [bs] * dim, g, neq, n, n_work_slots=0)
    f.attach(pool, depth, order)
    pool.upload(L.WH_BLOCK, 0, hvy)
    pool.sync(L.WH_BLOCK, 0)
    got = pool.download(L.WH_BLOCK, 0, n)
    want = hvy.copy()
    po.LeafForest(dim, jmax, [bs] * dim, g, tcs, lev).sync(want, depth, depth, order)
    assert np.array_equal(got, want)
    pool.close()
    f.close()


@pytest.mark.parametrize("wav,order,g", [("CDF20", 2, 1), ("CDF40", 4, 3), ("CDF60", 6, 5)])
def test_refine_everywhere_reference(wav, order, g):
    """wabbit-post --refine-everywhere on the device: sync_ghosts_tree,
    refineBlock of all 112 blocks, sync before saving -- equal to the
    reference's vor_00100.h5."""
    src = load_npz("wavelet_input.npz")
    bs, jmax = int(src["bs"][0]), int(src["max_level"])
    n = src["treecode"].shape[0]
    hvy = fixture_to_hvy([src["field_vor"]], src["origin"], src["spacing"], bs, g, 2)
    nb_max = 4 * n + 8
    f = Forest(2, jmax, [bs, bs], g, [float(src["domain_size"][0])] * 2, nb_max=nb_max)
    f.set_blocks(src["treecode"], src["level"], np.arange(n, dtype=np.int32))
    pool = L.Pool(2, [bs, bs], g, 1, nb_max, n_work_slots=0)
    f.attach(pool, g, order)
    pool.upload(L.WH_BLOCK, 0, hvy)
    refine_on_device(pool, f, np.arange(n, dtype=np.int32), order, g)
    pool.sync(L.WH_BLOCK, 1)
    tc, lv, hv, _ = f.blocks()
    allb = pool.download(L.WH_BLOCK, 0, nb_max)
    ref = load_npz(f"wavelet_refine_{wav}.npz")
    index = {(int(t), int(l)): int(h) for t, l, h in zip(tc, lv, hv)}
    assert len(index) == params_of(ref).get("n_total", 0) or len(index) == ref["treecode"].shape[0]
    mine = np.stack([allb[index[(int(t), int(l))], 0, 0, g:g + bs + 1, g:g + bs + 1]
                     for t, l in zip(ref["treecode"], ref["level"])])
    assert rel_l2(mine, ref["field_vor"]) <= 1e-14
    assert np.array_equal(mine, ref["field_vor"])
    pool.close()
    f.close()


def _oracle_rk_step_leaf(of, acm, hvy, x0dx, g, g_rhs, bs, dim, dt, bt):
    """RungeKuttaGeneric with the leaf-forest ghost fill (oracle)."""
    p = po.make_params(acm)
    n = hvy.shape[0]
    sl = (slice(None), slice(None)) + ((slice(g, g + bs),) * 3 if dim == 3 else (slice(None), slice(g, g + bs), slice(g, g + bs)))
    u = hvy.copy()
    of.sync(u, g_rhs, g_rhs, 4)
    u0 = u.copy()
    ks = []
    def rhs(v):
        return np.stack([po.rhs_block(p, [bs] * dim, g, x0dx[k, 3:3 + dim], v[k]) for k in range(n)])
    ks.append(rhs(u))
    s = bt.shape[0] - 1
    for j in range(2, s + 1):
        u = u.copy()
        u[sl] = u0[sl]
        for l in range(2, j + 1):
            a = bt[j - 1, l - 1]
            if abs(a) < 1e-8:
                continue
            u[sl] = u[sl] + (dt * a) * ks[l - 2][sl]
        of.sync(u, g_rhs, g_rhs, 4)
        ks.append(rhs(u))
    out = u.copy()
    out[sl] = u0[sl]
    for j in range(2, s + 2):
        b = bt[s, j - 1]
        if abs(b) < 1e-8:
            continue
        out[sl] = out[sl] + (dt * b) * ks[j - 2][sl]
    return out


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("path", [0, 2])
def test_rk4_on_nonuniform_forest(dim, path):
    bs = 16 if dim == 2 else 8
    g, g_rhs = 3, 2
    jmax = 4 if dim == 2 else 3
    tcs, lev = random_graded_forest(dim, 1 if dim == 3 else 2, jmax, 7, nref=5 if dim == 2 else 2)
    n = tcs.shape[0]
    f = Forest(dim, jmax, [bs] * dim, g, [1.0] * dim, nb_max=n)
    f.set_blocks(tcs, lev, np.arange(n, dtype=np.int32))
    _, _, hv, x0dx = f.blocks()
    rng = np.random.default_rng(2)
    hvy = 0.3 * random_block_data(n, dim + 1, dim, bs, g, 9)
    acm = ACMParams(dim=dim, scheme=4, skew_symmetry=False, penalization=False, c_0=10.0, nu=1e-2,
                    gamma_p=1.0, CFL=0.5, domain_size=[1.0] * 3)
    pool = L.Pool(dim, [bs] * dim, g, dim + 1, n, n_work_slots=5)
    f.attach(pool, g_rhs, 4)
    pool.upload(L.WH_BLOCK, 0, hvy)
    ac = L.acm_params_c(acm)
    dt = clamp_dt(pool.dt(ac), 0.0, TimeParams(time_max=10.0))
    pool.rk_step(ac, butcher_rk4(), 0.0, dt, path)
    got = pool.download(L.WH_BLOCK, 0, n)
    of = po.LeafForest(dim, jmax, [bs] * dim, g, tcs, lev)
    want = _oracle_rk_step_leaf(of, acm, hvy, x0dx, g, g_rhs, bs, dim, dt, butcher_rk4())
    gi = interior(got, bs, g, dim)
    wi = interior(want, bs, g, dim)
    assert rel_l2(gi, wi) <= 1e-12
    print("bit-identical:", np.array_equal(gi, wi))
    pool.close()
    f.close()


@pytest.mark.parametrize("order,g,bs", [(2, 2, 8), (4, 3, 12), (6, 5, 12), (4, 6, 24)])
def test_refine_blocks_3d_matches_oracle(order, g, bs):
    """refineBlock in 3D (wh_refine_blocks, the LDS plane-marching kernel)
    against the oracle's refineBlock (refinementExecute.f90:4-116): every
    daughter of every mother bit for bit, for the three predictor orders,
    daughters listed in a scrambled order of mothers."""
    import ctypes
    rng = np.random.default_rng(order * 100 + bs)
    nm, neq, nd = 3, 2, 8
    n = bs + 2 * g
    moms = rng.standard_normal((nm, neq, n, n, n))
    nb = nm * (nd + 1)
    pool = L.Pool(3, [bs] * 3, g, neq, nb, n_work_slots=0)
    pool.upload(L.WH_BLOCK, 0, moms)
    tri = []
    for m in (2, 0, 1):
        for k in (3, 0, 7, 1, 6, 2, 5, 4):
            tri.append((m, nm + m * nd + k, k))
    t = np.ascontiguousarray(np.array(tri, np.int32))
    L.check(L.load().wh_refine_blocks(pool.h, L.WH_BLOCK, t.ctypes.data, t.shape[0], order))
    got = pool.download(L.WH_BLOCK, 0, nb)
    pool.close()
    lib = po.lib()
    for m in range(nm):
        daughters = [np.zeros((neq, n, n, n)) for _ in range(nd)]
        arr = (ctypes.POINTER(ctypes.c_double) * nd)(
            *[d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for d in daughters])
        mm = np.ascontiguousarray(moms[m])
        bsa = np.array([bs] * 3, np.int32)
        lib.or_refine_block(3, bsa.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), g, neq,
                            mm.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), arr, order)
        for k in range(nd):
            assert np.array_equal(got[nm + m * nd + k], daughters[k]), (m, k)
