"""GPU: the 3D penalisation mask (create_mask_3D_acm "sphere-fixed" with
sponge_3D, create_mask.f90:6-174, sponge.f90:61-233) against the oracle, and
a penalised 3D adaptive run (sphere in a periodic cube with a rect sponge)
against the oracle's main loop.

Bars: chi, u_s and colour within 4e-16 (the device sqrt/cos against host
libm); the p-norm sponge within 1e-14 (device pow); the adaptive run with
identical block lists and fields within relative L2 1e-12."""
import numpy as np
import pytest

from helpers import random_graded_forest
from oracle import pyoracle as po
from wabbit_amd import lib as L
from wabbit_amd import wavelets as W
from wabbit_amd.adaptive import AdaptConfig, AdaptiveACM
from wabbit_amd.params import ACMParams, TimeParams, butcher_rk4, clamp_dt
from wabbit_amd.tree import Forest

pytestmark = pytest.mark.gpu


def _mask_kw(sponge_type, L_sponge=0.15):
    return dict(geometry=2, x_cntr=(0.43, 0.5, 0.57), R_cyl=0.2, smoothing_width=0.02, smoothing_safety=0.02,
                use_sponge=True, sponge_type=sponge_type, L_sponge=L_sponge, p_sponge=8.0, domain=(1.0, 1.0, 1.0))


@pytest.mark.parametrize("sponge_type", [0, 1])
def test_mask3d_kernel_matches_oracle(sponge_type):
    bs, g, jmax = 8, 3, 3
    tcs, lev = random_graded_forest(3, 1, jmax, 4, nref=4)
    n = len(tcs)
    f = Forest(3, jmax, [bs] * 3, g, [1.0] * 3, nb_max=n)
    f.set_blocks(tcs, lev, np.arange(n, dtype=np.int32))
    _, _, hv, x0dx = f.blocks()
    pool = L.Pool(3, [bs] * 3, g, 4, n, n_work_slots=0, n_mask=6, mask_per_block=True)
    kw = _mask_kw(sponge_type)
    W.create_mask(pool, W.mask_params(**kw), hv, x0dx)
    got = pool.download(L.WH_MASK, 0, n)
    mp = po.mask_params(**kw)
    want = np.stack([po.mask_block(mp, 3, [bs] * 3, g, x0dx[k, :3], x0dx[k, 3:]) for k in range(n)])
    assert (want[:, 0] > 0).any() and (want[:, 0] == 1.0).any() and (want[:, 5] > 0).any()
    d = np.abs(got[hv] - want)
    assert np.max(d[:, :5]) <= 4e-16
    assert np.max(d[:, 5]) <= (0.0 if sponge_type == 0 else 1e-14)
    pool.close()
    f.close()


def test_mask3d_geometry_checked():
    pool = L.Pool(3, [8] * 3, 3, 4, 2, n_work_slots=0, n_mask=6, mask_per_block=True)
    kw = _mask_kw(0)
    kw["geometry"] = 1                                   # the 2D cylinder is not a 3D geometry
    with pytest.raises(L.WabbitHipError):
        W.create_mask(pool, W.mask_params(**kw), np.zeros(1, np.int32), np.array([[0, 0, 0, .1, .1, .1]]))
    pool.close()


def _ic(tc, lv, jmax, bs, g):
    """u = (1, 0, 0) plus a small seeded smooth perturbation, p = 0."""
    from wabbit_amd.forest import ijk_from_treecode
    n1 = bs + 2 * g
    out = np.zeros((len(tc), 4, n1, n1, n1))
    for k, (t, l) in enumerate(zip(tc, lv)):
        ijk = ijk_from_treecode(int(t), int(l), jmax, 3)
        dx = 2.0 ** (-int(l)) / bs
        ax = [ijk[d] * bs * dx + (np.arange(n1) - g) * dx for d in range(3)]
        Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        out[k, 0] = 1.0 + 0.05 * np.sin(2 * np.pi * Y) * np.cos(2 * np.pi * Z)
        out[k, 1] = 0.05 * np.sin(2 * np.pi * X + 0.3)
        out[k, 2] = 0.05 * np.cos(2 * np.pi * (X + Y))
    return out


def test_penalised_3d_adaptive_run():
    # Bs = 16: the smallest CDF44 block the reconstruction needs no neighbours for
    # (Bs >= Nreconl/r, adapt_tree.f90:739); 14 s of oracle time for two steps
    bs, jmin, jmax, nsteps = 16, 1, 3, 2
    acm = ACMParams(dim=3, scheme=4, skew_symmetry=False, penalization=True, use_sponge=True, c_0=10.0, nu=1e-2,
                    gamma_p=1.0, C_eta=1e-2, C_sponge=1e-1, u_mean_set=[1.0, 0.0, 0.0], CFL=0.5,
                    domain_size=[1.0, 1.0, 1.0])
    cfg = AdaptConfig(wavelet="CDF44", bs=bs, jmin=jmin, jmax=jmax, eps=1e-2, groups=[1, 1, 1, 1])
    kw = _mask_kw(0)
    w = W.setup("CDF44", 2)
    from wabbit_amd.forest import equidistant_treecodes
    tc, lv = equidistant_treecodes(1, 3, jmax)
    u0 = _ic(tc, lv, jmax, bs, w.g)
    tp = TimeParams(time_max=1e9)
    s = AdaptiveACM(acm, tp, cfg, tc, lv, u0, mask=W.mask_params(**kw), nb_max=4096, jmax_treecode=jmax)
    # the oracle from the same leaves
    wo = po.wavelet("CDF44", fd_max_size=2)
    mp = po.mask_params(**kw)
    tree = po.Tree(wo, 3, jmin, jmax, [bs] * 3, wo.g, tc, lv, u0, indicator=1, eps=cfg.eps, groups=cfg.groups,
                   domain=(1.0, 1.0, 1.0), mask=mp)
    ap = po.make_params(acm)
    t = 0.0
    for _ in range(nsteps):
        s.step()
        tree.sync()
        tree.refine_everywhere()
        dt = clamp_dt(tree.dt(ap), t, tp)
        tree.rk_step(ap, mp, butcher_rk4(), dt, 2)
        t += dt
        tree.adapt()
    tcg, lvg, data = s.state()
    tco, lvo, want = tree.get()
    tree.close()
    s.close()
    assert abs(s.time - t) <= 1e-15 * t
    assert np.array_equal(tcg, tco) and np.array_equal(lvg, lvo)
    assert len(set(lvg.tolist())) >= 2                    # adapted to a mixed-level grid
    g = w.g
    inner = (slice(None), slice(None), slice(g, g + bs), slice(g, g + bs), slice(g, g + bs))
    err = np.linalg.norm(data[inner] - want[inner]) / np.linalg.norm(want[inner])
    print("3D penalised run:", len(tcg), "blocks, rel L2", err)
    assert err <= 1e-12
