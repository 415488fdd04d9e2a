"""The adaptive channel-cylinder workload (wabbit_amd.workloads.channel) on the
oracle: the initial grid of setInitialCondition_tree.f90:75-108 and the main
loop of LIB/MAIN/main.f90:161-230, at reduced Jmax for test sizes."""
import numpy as np

from oracle import pyoracle as po
from wabbit_amd import workloads as wl
from wabbit_amd.forest import equidistant_treecodes, refined_treecodes
from wabbit_amd.params import butcher_rk4, clamp_dt

G, GRHS = 6, 2


def oracle_tree(w, tc, lv, hvy):
    cfg = wl.channel_adapt_config(w)
    return po.Tree(po.wavelet(w["wavelet"], fd_max_size=2), 2, w["jmin"], w["jmax"], [w["bs"], w["bs"]], G,
                   tc, lv, hvy, indicator=1, eps=w["eps"], groups=cfg.groups,
                   force_maxlevel_dealiasing=cfg.force_maxlevel_dealiasing, eps_normalized=cfg.eps_normalized,
                   threshold_mask=cfg.threshold_mask, domain=w["domain"], mask=po.mask_params(**w["mask"]))


def oracle_initial_grid(w):
    """(treecodes, levels, data) of the initial grid, sorted by (level, treecode)."""
    ic = wl.ChannelInicond(w, G)
    jm = w["jmax"]
    tc, lv = equidistant_treecodes(w["jmin"], 2, jm)
    n_old, it = -1, 0
    while len(tc) != n_old and it < w["jmax"] - w["jmin"]:
        n_old = len(tc)
        tc, lv = refined_treecodes(tc, lv, jm, 2, jm)
        tree = oracle_tree(w, tc, lv, ic(tc, lv, jm))
        tree.adapt()
        tc, lv, hvy = tree.get()
        tree.close()
        it += 1
    if it == 0:
        hvy = ic(tc, lv, jm)
    return tc, lv, hvy


def oracle_run(w, tc, lv, hvy, nsteps, on_step=None):
    """nsteps main-loop iterations on the oracle; returns (tree, time)."""
    tree = oracle_tree(w, tc, lv, hvy)
    ap = po.make_params(w["acm"])
    mp = po.mask_params(**w["mask"])
    bt = butcher_rk4()
    time = 0.0
    for it in range(nsteps):
        tree.sync()
        tree.refine_everywhere()
        dt = clamp_dt(tree.dt(ap), time, w["tp"])
        tree.rk_step(ap, mp, bt, dt, GRHS)
        time += dt
        tree.adapt()
        if on_step:
            on_step(tree, time, it + 1)
    return tree, time


def oracle_steps(w, tc, lv, hvy):
    """The oracle's main loop as a generator of (time, treecodes, levels,
    hvy) after each iteration, so a test can compare step by step without
    holding every state."""
    tree = oracle_tree(w, tc, lv, hvy)
    ap = po.make_params(w["acm"])
    mp = po.mask_params(**w["mask"])
    bt = butcher_rk4()
    time = 0.0
    try:
        while True:
            tree.sync()
            tree.refine_everywhere()
            dt = clamp_dt(tree.dt(ap), time, w["tp"])
            tree.rk_step(ap, mp, bt, dt, GRHS)
            time += dt
            tree.adapt()
            yield (time,) + tuple(tree.get())
    finally:
        tree.close()


def sorted_keys(tc, lv):
    return sorted((int(a), int(b)) for a, b in zip(tc, lv))


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
