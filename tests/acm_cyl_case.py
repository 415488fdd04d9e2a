"""The adaptive penalised-cylinder case of the reference's TESTING/acm/acm_CDF44
(acm_cyl.ini) as parameters for the oracle and the device loop: the main
loop of LIB/MAIN/main.f90:161-230 with adapt_tree = 1, refinement indicator
"everywhere", N_dt_per_grid = 1."""
import numpy as np

from helpers import fixture_to_hvy, load_npz, params_of
from oracle import pyoracle as po
from wabbit_amd.params import ACMParams, TimeParams, butcher_rk4, clamp_dt

BS, G, GRHS, JMIN, JMAX = 26, 6, 2, 1, 6
L = 20.0


def acm_params():
    return ACMParams(dim=2, scheme=4, skew_symmetry=False, penalization=True, use_sponge=True,
                     c_0=12.5, nu=0.0, gamma_p=0.0, C_eta=1.34e-3, C_sponge=8.0e-3,
                     u_mean_set=[0.0, -1.0, 0.0], CFL=1.5, CFL_eta=0.99, domain_size=[L, L, 1.0])


def time_params():
    return TimeParams(time_max=0.1, write_method="fixed_time", write_time=0.05, tsave_stats=0.2)


def smoothing():
    """module_ACM.f90:460-471: cosine smoothing over C_smooth dx_min, dx_min
    of the finest level."""
    dx_min = 2.0 ** (-JMAX) * L / BS
    w = dx_min * 1.5
    return w, w


def oracle_mask_params():
    w, s = smoothing()
    return po.mask_params(geometry=1, x_cntr=(10.0, 10.0, 0.0), R_cyl=0.5, smoothing_width=w,
                          smoothing_safety=s, use_sponge=True, sponge_type=1, L_sponge=2.0, p_sponge=8.0,
                          domain=(L, L, 1.0))


def initial_state():
    f = load_npz("acm_cyl_t0.npz")
    hvy = fixture_to_hvy([f["field_ux"], f["field_uy"], f["field_p"]], f["origin"], f["spacing"], BS, G, 2)
    return f["treecode"], f["level"], hvy, int(f["max_level"])


def oracle_tree(tc, lv, hvy, jmax_tc):
    w = po.wavelet("CDF44", fd_max_size=2)
    return po.Tree(w, 2, JMIN, JMAX, [BS, BS], G, tc, lv, hvy, indicator=1, eps=1e-3, groups=[1, 1, 1],
                   force_maxlevel_dealiasing=True, eps_normalized=True, threshold_mask=True,
                   domain=(L, L, 1.0), mask=oracle_mask_params())


def oracle_run(t_end, on_step=None):
    """The reference's main loop on the oracle from the saved t=0 state to
    t_end; returns the tree and (time, iteration)."""
    tc, lv, hvy, jmax_tc = initial_state()
    assert jmax_tc == JMAX
    tree = oracle_tree(tc, lv, hvy, jmax_tc)
    acm = acm_params()
    ap = po.make_params(acm)
    mp = oracle_mask_params()
    tp = time_params()
    bt = butcher_rk4()
    time, it = 0.0, 0
    while time < t_end - 1e-12:
        tree.sync()
        tree.refine_everywhere()
        # RungeKuttaGeneric: sync_ghosts_RHS_tree then calculate_time_step
        dt = clamp_dt(tree.dt(ap), time, tp)
        tree.rk_step(ap, mp, bt, dt, GRHS)
        time += dt
        it += 1
        tree.adapt()
        if on_step:
            on_step(tree, time, it)
    return tree, time, it
