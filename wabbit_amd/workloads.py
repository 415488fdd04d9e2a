"""The benchmark workloads: parameters, seeded inputs and initial grids.

c1   2D ACM, periodic unit square, uniform level 4 (256 blocks), Bs=16,
     CDF40 ghosts (g=3, g_rhs=2), FD_4th_central, RK4, CFL 0.5, c0=20,
     nu=1e-3; seeded Fourier velocity (|k|<=8, amplitude |k|^-2, u_rms=1)
     and a Gaussian pressure pulse (sigma 0.05, amplitude 0.1).
c2   2D ACM, adaptive, penalised cylinder in a channel: domain [0,4]x[0,2]
     with a rectangular sponge layer (the periodic domain's inflow/outflow),
     Bs=16, Jmin=2, Jmax=6, CDF44 (g=6, g_rhs=2), eps=1e-3 (absolute),
     adapt every step.  Cylinder of seeded radius in
     [0.1,0.2] and seeded centre, C_eta=1e-3, mask smoothed over 1.5 dx_min.
     Inflow u=(1,0) plus a seeded Fourier perturbation (|k|<=8, |k|^-2,
     u_rms=0.05), p=0; nu=1e-3, c0=20, CFL 0.5 (the reference's default,
     ini_file_to_params.f90:624), RK4, 500 steps.
c3u  3D ACM, periodic unit cube, uniform level 4 (4096 blocks), Bs=24,
     CDF44 ghosts (g=6, g_rhs=2), FD_4th_central, c0=15, nu=1e-3; seeded
     solenoidal velocity (1<=|k|<=6, E(k)~k^-5/3, u_rms=1), p=0.

The adaptive workload builds its initial grid the way the reference's
setInitialCondition_tree does for an analytic initial condition
(LIB/MESH/setInitialCondition_tree.f90:75-108): an equidistant grid at
Jini=Jmin, then (refine everywhere, set the initial condition on every
block, adapt_tree) until the block count stops changing or Jmax-Jmin
rounds are done.
"""
from __future__ import annotations

import numpy as np

from .forest import ijk_from_treecode
from .inicond import fourier_velocity_grid
from .params import ACMParams, TimeParams


def get(name: str) -> dict:
    if name == "c1":
        return dict(name="c1-2d-periodic-J4-Bs16", kind="uniform", dim=2, J=4, bs=16, g=3, g_rhs=2,
                    L=1.0,
                    acm=ACMParams(dim=2, scheme=4, skew_symmetry=False, penalization=False,
                                  c_0=20.0, nu=1e-3, gamma_p=1.0, CFL=0.5,
                                  domain_size=[1.0, 1.0, 1.0]),
                    ic=dict(kmax=8, kmin=1, amp_exponent=-2.0), pressure=True, nsteps_cfg=100)
    if name == "c2":
        return channel()
    if name == "c3":
        return periodic3d()
    if name == "c3u":
        return dict(name="c3u-3d-periodic-J4-Bs24", kind="uniform", dim=3, J=4, bs=24, g=6, g_rhs=2,
                    L=1.0,
                    acm=ACMParams(dim=3, scheme=4, skew_symmetry=False, penalization=False,
                                  c_0=15.0, nu=1e-3, gamma_p=1.0, CFL=0.5,
                                  domain_size=[1.0, 1.0, 1.0]),
                    ic=dict(kmax=6, kmin=1, energy_exponent=-5.0 / 3.0, solenoidal=True),
                    pressure=False, nsteps_cfg=200)
    raise ValueError(f"unknown workload {name}")


def channel(seed: int = 1234, jmax: int = 6, bs: int = 16) -> dict:
    """The adaptive penalised-cylinder channel (c2) with its seeded geometry."""
    rng = np.random.default_rng(seed + 104729)
    domain = (4.0, 2.0, 1.0)
    R = float(rng.uniform(0.1, 0.2))
    xc = (float(rng.uniform(0.75, 1.25)), float(rng.uniform(0.75, 1.25)), 0.0)
    jmin = 2
    # module_ACM.f90:460-471: smoothing over C_smooth = 1.5 finest spacings
    dx_min = 2.0 ** (-jmax) * min(domain[0], domain[1]) / bs
    smooth = 1.5 * dx_min
    mask = dict(geometry=1, x_cntr=xc, R_cyl=R, smoothing_width=smooth, smoothing_safety=smooth,
                use_sponge=True, sponge_type=0, L_sponge=0.25, p_sponge=20.0, domain=domain)
    acm = ACMParams(dim=2, scheme=4, skew_symmetry=False, penalization=True, use_sponge=True,
                    c_0=20.0, nu=1e-3, gamma_p=1.0, C_eta=1e-3, C_sponge=1e-2,
                    u_mean_set=[1.0, 0.0, 0.0], CFL=0.5, CFL_eta=0.99,
                    domain_size=list(domain))
    return dict(name=f"c2-2d-channel-cylinder-J{jmin}-{jmax}-Bs{bs}", kind="adaptive", dim=2, bs=bs,
                jmin=jmin, jmax=jmax, wavelet="CDF44", eps=1e-3, domain=domain, acm=acm,
                tp=TimeParams(time_max=1e9), mask=mask, seed=seed,
                perturb=dict(kmax=8, kmin=1, amp_exponent=-2.0, urms=0.05), inflow=(1.0, 0.0),
                nsteps_cfg=500)


def sample_tree(treecodes, levels, jmax_tc, dim, bs, g, grid, level_fine):
    """Sample a global periodic grid (axes z,y,x; 2^level_fine*Bs points per
    direction) into the listed blocks incl. ghosts: (n, nz, ny, nx).  A block
    at level J takes every 2^(level_fine-J)-th point."""
    N = grid.shape[-1]
    nx, ny = bs + 2 * g, bs + 2 * g
    nz = bs + 2 * g if dim == 3 else 1
    out = np.empty((len(treecodes), nz, ny, nx))
    ar = np.arange(nx) - g
    for k, (tc, lv) in enumerate(zip(treecodes, levels)):
        ijk = ijk_from_treecode(int(tc), int(lv), jmax_tc, dim)
        st = 2 ** (level_fine - int(lv))
        ix = ((ijk[0] * bs + ar) * st) % N
        iy = ((ijk[1] * bs + ar) * st) % N
        if dim == 2:
            out[k, 0] = grid[np.ix_(iy, ix)]
        else:
            iz = ((ijk[2] * bs + ar) * st) % N
            out[k] = grid[np.ix_(iz, iy, ix)]
    return out


class ChannelInicond:
    """The c2 initial condition as a function of the block list: u = inflow +
    perturbation sampled from the finest-level grid, p = 0 (the analytic
    setInicondBlocks of the workload)."""

    def __init__(self, w: dict, g: int):
        self.w, self.g = w, g
        self.comps = None

    def __call__(self, treecodes, levels, jmax_tc):
        w = self.w
        if self.comps is None:
            N = 2 ** w["jmax"] * w["bs"]
            self.comps = fourier_velocity_grid(2, N, w["seed"], **w["perturb"])
        u = [sample_tree(treecodes, levels, jmax_tc, 2, w["bs"], self.g, c, w["jmax"]) + w["inflow"][d]
             for d, c in enumerate(self.comps)]
        return np.stack(u + [np.zeros_like(u[0])], axis=1)



def channel_adapt_config(w: dict):
    """[Blocks] settings of the channel workload, the reference's defaults
    (ini_file_to_params.f90:470-560): threshold indicator on all components
    with an absolute eps (eps_normalized off), no mask criterion, CDF44 with
    its lifted defaults (coarse extension, security zone)."""
    from .adaptive import AdaptConfig
    return AdaptConfig(wavelet=w["wavelet"], bs=w["bs"], jmin=w["jmin"], jmax=w["jmax"], eps=w["eps"],
                       eps_normalized=False, groups=[1, 1, 1], threshold_mask=False,
                       force_maxlevel_dealiasing=False)


def channel_solver(w: dict, path=None, nb_max: int = 16384):
    """The device solver of the channel workload on its initial grid."""
    from . import wavelets as W
    from .adaptive import AdaptiveACM, fd_reach
    from .forest import equidistant_treecodes
    from .timestep import PATH_FUSED
    cfg = channel_adapt_config(w)
    g = W.setup(w["wavelet"], fd_reach(w["acm"].scheme)).g
    ic = ChannelInicond(w, g)
    tc, lv = equidistant_treecodes(w["jmin"], 2, w["jmax"])
    s = AdaptiveACM(w["acm"], w["tp"], cfg, tc, lv, ic(tc, lv, w["jmax"]), mask=W.mask_params(**w["mask"]),
                    nb_max=nb_max, path=PATH_FUSED if path is None else path, jmax_treecode=w["jmax"])
    s.set_initial_grid(ic)
    return s


def channel_spec(seed: int, jmax: int, bs: int, nsteps: int, slots: int, tag: str = ""):
    """The channel workload with its seeded cylinder as one solve of a batch."""
    from . import wavelets as W
    from .adaptive import fd_reach
    from .batch import SolveSpec
    w = channel(seed=seed, jmax=jmax, bs=bs)
    g = W.setup(w["wavelet"], fd_reach(w["acm"].scheme)).g
    return SolveSpec(acm=w["acm"], cfg=channel_adapt_config(w), inicond=ChannelInicond(w, g), nsteps=nsteps,
                     mask=w["mask"], tp=w["tp"], slots=slots, tag=tag or w["name"] + f"-seed{seed}")


def ensemble(n: int = 64, seed: int = 4321, nsteps: int = 200):
    """The batched ensemble: n channel solves, each with its own seeded
    cylinder, Jmax drawn from {4,5,6,7} and Bs from {16,24}; slots per solve
    sized from Jmax."""
    rng = np.random.default_rng(seed)
    slots_of = {4: 1024, 5: 2560, 6: 6144, 7: 12288}
    out = []
    for i in range(n):
        jmax = int(rng.choice([4, 5, 6, 7]))
        bs = int(rng.choice([16, 24]))
        out.append(channel_spec(seed + 17 * i, jmax, bs, nsteps, slots_of[jmax], tag=f"member{i}-J{jmax}-Bs{bs}"))
    return out


def periodic3d(seed: int = 1234, jmax: int = 5, bs: int = 24, jmin: int = 2) -> dict:
    """The adaptive 3D periodic workload (c3): unit cube, Bs=24, Jmin=2,
    Jmax=5, CDF44, eps=1e-3; seeded solenoidal velocity (projected Fourier
    modes 1<=|k|<=6 with a k^-5/3 energy envelope, u_rms=1), p=0; nu=1e-3,
    c0=15, CFL 0.5, 200 RK4 steps."""
    acm = ACMParams(dim=3, scheme=4, skew_symmetry=False, penalization=False, use_sponge=False,
                    c_0=15.0, nu=1e-3, gamma_p=1.0, CFL=0.5, domain_size=[1.0, 1.0, 1.0])
    return dict(name=f"c3-3d-periodic-J{jmin}-{jmax}-Bs{bs}", kind="adaptive3d", dim=3, bs=bs, jmin=jmin,
                jmax=jmax, wavelet="CDF44", eps=1e-3, domain=(1.0, 1.0, 1.0), acm=acm,
                tp=TimeParams(time_max=1e9), mask=None, seed=seed,
                modes=dict(kmax=6, kmin=1, energy_exponent=-5.0 / 3.0, solenoidal=True), nsteps_cfg=200)


class PeriodicInicond:
    """Seeded Fourier velocity and p = 0 as a function of the block list.
    The field of a block at level J is sampled from the periodic grid of
    level J (2^J Bs points per direction), built once per level on first
    use: the modes (and the u_rms normalisation, exact on any grid finer
    than the highest mode) are the same on every level."""

    def __init__(self, w: dict, g: int):
        self.w, self.g = w, g
        self.grids = {}

    def grid(self, level):
        if level not in self.grids:
            w = self.w
            N = 2 ** level * w["bs"]
            self.grids[level] = fourier_velocity_grid(w["dim"], N, w["seed"], **w["modes"])
        return self.grids[level]

    def __call__(self, treecodes, levels, jmax_tc):
        w, dim = self.w, self.w["dim"]
        tc = np.asarray(treecodes)
        lv = np.asarray(levels)
        nz = w["bs"] + 2 * self.g if dim == 3 else 1
        n1 = w["bs"] + 2 * self.g
        out = np.zeros((len(tc), dim + 1, nz, n1, n1))
        for L in sorted(set(int(x) for x in lv)):
            sel = np.nonzero(lv == L)[0]
            comps = self.grid(L)
            for d in range(dim):
                out[sel, d] = sample_tree(tc[sel], lv[sel], jmax_tc, dim, w["bs"], self.g, comps[d], L)
        return out


def adaptive_solver(w: dict, path=None, nb_max: int = 16384):
    """The device solver of an adaptive workload (c2 or c3) on its initial grid."""
    if w["kind"] == "adaptive":
        return channel_solver(w, path=path, nb_max=nb_max)
    from . import wavelets as W
    from .adaptive import AdaptConfig, AdaptiveACM, fd_reach
    from .forest import equidistant_treecodes
    from .timestep import PATH_FUSED
    cfg = AdaptConfig(wavelet=w["wavelet"], bs=w["bs"], jmin=w["jmin"], jmax=w["jmax"], eps=w["eps"],
                      groups=[1] * (w["dim"] + 1))
    g = W.setup(w["wavelet"], fd_reach(w["acm"].scheme)).g
    ic = PeriodicInicond(w, g)
    tc, lv = equidistant_treecodes(w["jmin"], w["dim"], w["jmax"])
    s = AdaptiveACM(w["acm"], w["tp"], cfg, tc, lv, ic(tc, lv, w["jmax"]), mask=None, nb_max=nb_max,
                    path=PATH_FUSED if path is None else path, jmax_treecode=w["jmax"])
    s.set_initial_grid(ic)
    return s


def periodic3d_spec(seed: int, jmax: int, bs: int, nsteps: int, slots: int, tag: str = ""):
    """The 3D periodic workload with its own seeded field as one solve of a batch."""
    from . import wavelets as W
    from .adaptive import AdaptConfig, fd_reach
    from .batch import SolveSpec
    w = periodic3d(seed=seed, jmax=jmax, bs=bs)
    g = W.setup(w["wavelet"], fd_reach(w["acm"].scheme)).g
    cfg = AdaptConfig(wavelet=w["wavelet"], bs=bs, jmin=w["jmin"], jmax=jmax, eps=w["eps"], groups=[1, 1, 1, 1])
    return SolveSpec(acm=w["acm"], cfg=cfg, inicond=PeriodicInicond(w, g), nsteps=nsteps, mask=None, tp=w["tp"],
                     slots=slots, tag=tag or w["name"] + f"-seed{seed}")


# pool slots reserved per solve, by dimension and Jmax
SLOTS_2D = {4: 1024, 5: 2560, 6: 6144, 7: 12288}
SLOTS_3D = {3: 768, 4: 2048, 5: 4096}


def serving_mix(n: int = 256, seed: int = 2024, rate: float = 4.0, frac_2d: float = 0.7):
    """The serving stream: n requests with Poisson arrivals of rate `rate`
    per second; 70% 2D penalised channel solves (Jmax 4-7, Bs 16/24, 100-500
    steps), 30% 3D periodic solves (Jmax 4-5, Bs 24, 50-200 steps).  Returns
    [(rid, spec, arrival seconds)]."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        s = seed * 7919 + 13 * i
        if rng.random() < frac_2d:
            jmax, bs = int(rng.choice([4, 5, 6, 7])), int(rng.choice([16, 24]))
            steps = int(rng.integers(100, 501))
            spec = channel_spec(s, jmax, bs, steps, SLOTS_2D[jmax], tag=f"req{i}-2d-J{jmax}-Bs{bs}-{steps}st")
        else:
            jmax = int(rng.choice([4, 5]))
            steps = int(rng.integers(50, 201))
            spec = periodic3d_spec(s, jmax, 24, steps, SLOTS_3D[jmax], tag=f"req{i}-3d-J{jmax}-Bs24-{steps}st")
        out.append((i, spec, t))
    return out
