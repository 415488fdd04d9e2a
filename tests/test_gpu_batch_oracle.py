"""GPU: the batched path (one fused stage launch per RK stage over all pools
of a dimension, wh_batch_rk_step) against the oracle and against itself.

  * a mixed batch -- Bs 16 and 24, Jmax 4 and 5, two host lanes, so four
    pools in one launch -- with every member compared to its own oracle run
    of the reference's main loop (sync_ghosts_tree, refine everywhere, mask,
    RK4 with the host dt clamps, adapt_tree): block lists identical, fields
    within relative L2 1e-12 (the device mask's cos may differ from the host
    libm by an ulp);
  * physics constants per group: members that differ in c_0, nu, C_eta and
    the inflow share one pool and each equals its run alone, bit for bit;
  * the step replayed from its hipGraph equals the step launched eagerly,
    over several steps whose grids change, bit for bit.
"""
import dataclasses

import numpy as np
import pytest

from oracle import pyoracle as po
from wabbit_amd import workloads as wl
from wabbit_amd.batch import Batch, run_alone
from wabbit_amd.params import butcher_rk4, clamp_dt

pytestmark = pytest.mark.gpu

NSTEPS = 3


def _oracle_run(spec, state, nsteps):
    """The oracle's main loop from a member's initial grid."""
    tc, lv, hvy = state
    cfg = spec.cfg
    wo = po.wavelet(cfg.wavelet, fd_max_size=2)
    mp = po.mask_params(**spec.mask)
    dom = spec.acm.domain_size
    tree = po.Tree(wo, 2, cfg.jmin, cfg.jmax, [cfg.bs] * 2, wo.g, tc, lv, hvy, indicator=1, eps=cfg.eps,
                   groups=cfg.groups, eps_normalized=cfg.eps_normalized, threshold_mask=cfg.threshold_mask,
                   domain=dom, mask=mp)
    ap = po.make_params(spec.acm)
    t = 0.0
    for _ in range(nsteps):
        tree.sync()
        tree.refine_everywhere()
        dt = clamp_dt(tree.dt(ap), t, spec.tp)
        tree.rk_step(ap, mp, butcher_rk4(), dt, 2)
        t += dt
        tree.adapt()
    out = tree.get()
    tree.close()
    return out, t


def test_mixed_batch_members_match_oracle():
    specs = [wl.channel_spec(41, 4, 16, NSTEPS, 1024), wl.channel_spec(42, 5, 16, NSTEPS, 2560),
             wl.channel_spec(43, 4, 24, NSTEPS, 1024), wl.channel_spec(44, 5, 24, NSTEPS, 2560)]
    b = Batch({(2, 16): 2 * 4096, (2, 24): 2 * 4096}, lanes=2)
    try:
        solves = [b.admit(x) for x in specs]
        assert all(s is not None for s in solves)
        assert len(b.groups) == 4                        # 2 lanes x 2 block sizes, one launch per stage
        init = [s.state() for s in solves]
        while not all(s.done for s in solves):
            b.step()
        got = [(s.state(), s.time) for s in solves]
    finally:
        b.close()
    for spec, st0, ((tc, lv, data), t) in zip(specs, init, got):
        (tco, lvo, want), to = _oracle_run(spec, st0, NSTEPS)
        assert np.array_equal(tc, tco) and np.array_equal(lv, lvo), spec.tag
        assert t == to
        g, bs = (data.shape[-1] - spec.cfg.bs) // 2, spec.cfg.bs
        inner = (slice(None), slice(None), 0, slice(g, g + bs), slice(g, g + bs))
        err = np.linalg.norm(data[inner] - want[inner]) / np.linalg.norm(want[inner])
        print(spec.tag, len(tc), "blocks, rel L2", err, flush=True)
        assert err <= 1e-12, spec.tag


def _variant(seed, jmax, c0, nu, C_eta, u_in):
    sp = wl.channel_spec(seed, jmax, 16, NSTEPS, 1024 if jmax == 4 else 2560)
    acm = dataclasses.replace(sp.acm, c_0=c0, nu=nu, C_eta=C_eta, u_mean_set=[u_in, 0.0, 0.0])
    return dataclasses.replace(sp, acm=acm)


def test_group_constants_in_one_pool():
    specs = [_variant(51, 4, 20.0, 1e-3, 1e-3, 1.0), _variant(52, 4, 12.0, 5e-3, 2e-3, 0.8),
             _variant(53, 5, 25.0, 2e-4, 5e-4, 1.2)]
    alone = [run_alone(x, 4096) for x in specs]
    b = Batch({(2, 16): 8192})
    try:
        solves = [b.admit(x) for x in specs]
        assert len(b.groups) == 1
        while not all(s.done for s in solves):
            b.step()
        mixed = [s.state() for s in solves]
    finally:
        b.close()
    for a, m, sp in zip(alone, mixed, specs):
        assert all(np.array_equal(x, y) for x, y in zip(a, m)), sp.tag
    # the constants matter: the members' own runs differ from one with the first one's constants
    other = run_alone(dataclasses.replace(specs[1], acm=specs[0].acm), 4096)
    assert not all(np.array_equal(x, y) for x, y in zip(other, alone[1]))


def test_graph_replay_matches_eager():
    specs = [wl.channel_spec(61, 4, 16, 5, 1024), wl.channel_spec(62, 4, 24, 5, 1024)]
    res = {}
    for g in (False, True):
        b = Batch({(2, 16): 4096, (2, 24): 4096}, use_graph=g)
        try:
            solves = [b.admit(x) for x in specs]
            while not all(s.done for s in solves):
                b.step()
            res[g] = [s.state() for s in solves]
        finally:
            b.close()
    for a, m in zip(res[False], res[True]):
        assert all(np.array_equal(x, y) for x, y in zip(a, m))
