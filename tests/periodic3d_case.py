"""The adaptive 3D periodic workload (wabbit_amd.workloads.periodic3d) on the
oracle: the initial grid of setInitialCondition_tree.f90:75-108 and the main
loop of LIB/MAIN/main.f90:161-230, at reduced Jmax and Bs for test sizes."""
import numpy as np

from oracle import pyoracle as po
from wabbit_amd import workloads as wl
from wabbit_amd.forest import equidistant_treecodes, refined_treecodes
from wabbit_amd.params import butcher_rk4, clamp_dt

G, GRHS = 6, 2


def oracle_tree(w, tc, lv, hvy):
    dim = w["dim"]
    return po.Tree(po.wavelet(w["wavelet"], fd_max_size=2), dim, w["jmin"], w["jmax"], [w["bs"]] * dim, G,
                   tc, lv, hvy, indicator=1, eps=w["eps"], groups=[1] * (dim + 1), domain=w["domain"])


def oracle_initial_grid(w):
    ic = wl.PeriodicInicond(w, G)
    jm, dim = w["jmax"], w["dim"]
    tc, lv = equidistant_treecodes(w["jmin"], dim, jm)
    hvy = ic(tc, lv, jm)
    n_old, it = -1, 0
    while len(tc) != n_old and it < w["jmax"] - w["jmin"]:
        n_old = len(tc)
        tc, lv = refined_treecodes(tc, lv, jm, dim, jm)
        tree = oracle_tree(w, tc, lv, ic(tc, lv, jm))
        tree.adapt()
        tc, lv, hvy = tree.get()
        tree.close()
        it += 1
    return tc, lv, hvy


def oracle_run(w, tc, lv, hvy, nsteps, on_step=None):
    tree = oracle_tree(w, tc, lv, hvy)
    ap = po.make_params(w["acm"])
    bt = butcher_rk4()
    time = 0.0
    for it in range(nsteps):
        tree.sync()
        tree.refine_everywhere()
        dt = clamp_dt(tree.dt(ap), time, w["tp"])
        tree.rk_step(ap, None, bt, dt, GRHS)
        time += dt
        tree.adapt()
        if on_step:
            on_step(tree, time, it + 1)
    return tree, time
