"""The adaptive ACM main loop on the device (LIB/MAIN/main.f90:161-230 with
adapt_tree = 1): per iteration sync_ghosts_tree, refine_tree "everywhere",
one RungeKuttaGeneric step (the mask rebuilt on the refined grid,
calculate_time_step with the host clamps), then adapt_tree with the
threshold indicator.  The forest metadata stays on the host (wh_forest_*),
all block data on the device; there is no CPU fallback.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import lib as L
from . import wavelets as W
from .params import ACMParams, TimeParams, butcher_rk4, clamp_dt
from .timestep import PATH_FUSED
from .forest import equidistant_treecodes, refined_treecodes
from .tree import Forest, refine_on_device


@dataclass
class AdaptConfig:
    """[Blocks] adaptivity keys (ini_file_to_params.f90:470-560)."""
    wavelet: str = "CDF44"
    bs: int = 16
    jmin: int = 1
    jmax: int = 6
    eps: float = 1e-3
    eps_normalized: bool = False
    groups: List[int] = field(default_factory=lambda: [1, 1, 1])
    threshold_mask: bool = False
    force_maxlevel_dealiasing: bool = False
    use_ce: Optional[bool] = None          # default: lifted wavelet
    security_zone: Optional[bool] = None   # default: lifted wavelet
    refinement_indicator: str = "everywhere"
    decompose_branch: int = 0              # 0 the reference's rule, 1 leaf-first, 2 level-wise
    eps_norm: int = 0                      # WH_NORM_*: params%eps_norm (Linfty by default)


def fd_reach(scheme: int) -> int:
    """FD_max_size of setup_wavelet: half-width of the FD stencils."""
    return {2: 1, 4: 2, 6: 3, 40: 3}[scheme]


class AdaptiveACM:
    """One adaptive ACM forest on the GPU."""

    def __init__(self, acm: ACMParams, tp: TimeParams, cfg: AdaptConfig, treecodes, levels, data,
                 mask: Optional[W.MaskParamsC] = None, nb_max: Optional[int] = None, path: int = PATH_FUSED,
                 jmax_treecode: Optional[int] = None):
        self.acm, self.tp, self.cfg, self.path = acm, tp, cfg, path
        self.w = W.setup(cfg.wavelet, fd_reach(acm.scheme))
        self.g = self.w.g
        self.g_rhs = self.w.g_rhs
        dim = acm.dim
        self.dim = dim
        n = len(treecodes)
        data = np.ascontiguousarray(data, np.float64)
        assert data.shape[1] == acm.neq
        self.nb_max = nb_max or max(8 * n, 256)
        self.mask_p = mask
        bs = [cfg.bs] * dim
        jtc = jmax_treecode if jmax_treecode is not None else cfg.jmax
        self.forest = Forest(dim, jtc, bs, self.g, acm.domain_size[:dim], self.nb_max)
        self.forest.set_blocks(treecodes, levels, np.arange(n, dtype=np.int32))
        use_mask = acm.penalization or acm.use_sponge
        # the fused RK path and adapt_tree need 3 work slots, the others one per RK4 stage
        self.pool = L.Pool(dim, bs, self.g, acm.neq, self.nb_max, n_work_slots=3 if path == PATH_FUSED else 5,
                           n_tmp_eqn=acm.neq, n_mask=6 if use_mask else 0, mask_per_block=True)
        W.set_wavelet(self.pool, self.w)
        self.pool.upload(L.WH_BLOCK, 0, data)
        lifted = bool(self.w.lifted)
        self.adapt_p = W.adapt_params(
            cfg.jmin, cfg.jmax, W.WH_IND_THRESHOLD,
            use_ce=lifted if cfg.use_ce is None else cfg.use_ce, eps=cfg.eps, groups=cfg.groups,
            force_maxlevel_dealiasing=cfg.force_maxlevel_dealiasing, eps_normalized=cfg.eps_normalized,
            threshold_mask=cfg.threshold_mask and mask is not None,
            security_zone=lifted if cfg.security_zone is None else cfg.security_zone,
            domain=acm.domain_size, mask=mask, decompose_branch=cfg.decompose_branch, eps_norm=cfg.eps_norm)
        self.acm_c = L.acm_params_c(acm)
        self.butcher = butcher_rk4()
        self.tp_c = L.time_params_c(tp)
        self.time, self.iteration = 0.0, 0
        # the fused path keeps dt, its clamps and the time advance on the
        # device (wh_run_steps) and learns them after adapt_tree without a
        # synchronisation of its own; (time, iteration) last set on the device
        self.device_time = path == PATH_FUSED
        self._dev_time = None
        self.rk_blocks = 0
        self.phase_ms = None
        self._stale = True
        self._attach()

    def _attach(self):
        self.forest.attach(self.pool, self.g_rhs, self.w.order, full=False)
        if self.mask_p is not None:
            W.create_mask(self.pool, self.mask_p)
        self._stale = False

    def _fresh(self):
        """The pool's tables and mask describe the current leaves."""
        if self._stale:
            self._attach()

    @property
    def nblocks(self):
        return self.forest.n

    def refine(self):
        """refine_tree "everywhere" (refine_tree.f90): every leaf below Jmax
        (respectJmaxJmin), on the ghost nodes synced by the caller."""
        tc, lv, hv, _ = self.forest.blocks()
        ids = hv[lv < self.cfg.jmax]
        if ids.size:
            refine_on_device(self.pool, self.forest, ids, self.w.order, self.g_rhs, presync=None, full=False)
        else:
            self.forest.attach(self.pool, self.g_rhs, self.w.order, full=False)
        self._stale = False

    def set_blocks(self, treecodes, levels, data):
        """Replace the leaves and their data (hvy ids 0..n-1), as
        setInicondBlocks_tree does on a new block list."""
        n = len(treecodes)
        if n > self.nb_max:
            raise ValueError(f"{n} blocks exceed the pool's {self.nb_max} slots")
        self.forest.set_blocks(treecodes, levels, np.arange(n, dtype=np.int32))
        self.pool.upload(L.WH_BLOCK, 0, np.ascontiguousarray(data, np.float64))
        self._attach()

    def adapt(self):
        """adapt_tree with the threshold indicator on the current leaves."""
        W.adapt_tree(self.pool, self.forest, self.adapt_p)
        self._attach()

    def set_initial_grid(self, inicond, jini: Optional[int] = None) -> int:
        """setInitialCondition_tree for an analytic initial condition
        (LIB/MESH/setInitialCondition_tree.f90:75-108): equidistant grid at
        Jini, then refine everywhere + set the initial condition on every
        block + adapt_tree until the block count stops changing or Jmax-Jmin
        rounds are done.  ``inicond(treecodes, levels, jmax_tc)`` returns
        the blocks' data (n, neq, nz, ny, nx).  Returns the number of rounds."""
        cfg, dim = self.cfg, self.dim
        jtc = self.forest.jmax
        tc, lv = equidistant_treecodes(cfg.jmin if jini is None else jini, dim, jtc)
        self.set_blocks(tc, lv, inicond(tc, lv, jtc))
        n_old, it = -1, 0
        while self.nblocks != n_old and it < cfg.jmax - cfg.jmin:
            n_old = self.nblocks
            tc, lv, _, _ = self.forest.blocks()
            tc, lv = refined_treecodes(tc, lv, jtc, dim, cfg.jmax)
            self.set_blocks(tc, lv, inicond(tc, lv, jtc))
            self.adapt()
            it += 1
        return it

    def autotune_next_step(self, regrow: float = 2.0):
        """Pick the stage-kernel variant on the next step's refined grid
        (the live blocks of a real RK stage; all variants are bit-identical),
        and again whenever the refined grid has grown by the factor
        ``regrow`` since the last pick (the fastest variant depends on the
        number of blocks per launch)."""
        self._tune = True
        self._regrow = regrow

    def step(self):
        """One main-loop iteration; returns dt.  ``rk_blocks`` holds the
        number of blocks the RK stages ran on (the refined grid)."""
        t = self._tick(None)
        W.sync_tree(self.pool, self.forest)
        t = self._tick(t, "sync_tree")
        self.refine()
        self.rk_blocks = self.forest.n
        tuned_at = getattr(self, "_tuned_at", None)
        regrown = (tuned_at and self.rk_blocks >= self._regrow * tuned_at
                   and not getattr(self.pool, "timing_on", False))   # no re-tuning while kernels are timed
        if getattr(self, "_tune", False) or regrown:
            self._tune = False
            self.tuned = self.pool.autotune(self.acm_c, reps=3)
            self._tuned_at = self.rk_blocks
        t = self._tick(t, "refine")
        if self.mask_p is not None:
            W.create_mask(self.pool, self.mask_p)
        t = self._tick(t, "mask")
        if self.device_time:
            if self._dev_time != (self.time, self.iteration):
                self.pool.time_set(self.time, self.iteration)
            self.pool.run_steps(self.acm_c, self.butcher, self.tp_c, 1, use_graph=False)
            self.pool.time_fetch()
        else:
            dt = clamp_dt(self.pool.dt(self.acm_c), self.time, self.tp)
            self.pool.rk_step(self.acm_c, self.butcher, self.time, dt, self.path)
            self.time += dt
            self.iteration += 1
        t = self._tick(t, "rk")
        W.adapt_tree(self.pool, self.forest, self.adapt_p)
        self._stale = True   # the next refine re-attaches; state()/mask() refresh
        if self.device_time:
            self.time, self.iteration, dt = self.pool.time_fetched()
            self._dev_time = (self.time, self.iteration)
        self._tick(t, "adapt")
        return dt

    def profile(self, on: bool):
        """Per-phase wall times (ms, accumulated in ``phase_ms``), with a
        stream synchronisation after each phase while on."""
        self.phase_ms = {} if on else None

    def _tick(self, t0, name=None):
        import time as _t
        if self.phase_ms is None:
            return None
        self.pool.synchronize()
        now = _t.perf_counter()
        if name is not None:
            self.phase_ms[name] = self.phase_ms.get(name, 0.0) + (now - t0) * 1e3
        return now

    def state(self):
        """(treecodes, levels, data) of the leaves, sorted by (level, treecode)."""
        tc, lv, hv, _ = self.forest.blocks()
        order = np.lexsort((tc, lv))
        data = np.stack([self.pool.download(L.WH_BLOCK, int(h), 1)[0] for h in hv[order]])
        return tc[order], lv[order], data

    def save(self, directory: str, names=None) -> list:
        """save_data (LIB/IO/save_data.f90:70-84): one WABBIT HDF5 file per
        component, ``<name>_<timestr>.h5``, written after a sync_ghosts_tree
        as saveHDF5_tree does (no_sync = .false.).  Returns the paths."""
        import os
        from . import h5io
        if names is None:
            names = ["ux", "uy", "uz"][:self.dim] + ["p"]
        W.sync_tree(self.pool, self.forest)
        tc, lv, data = self.state()
        files = h5io.files_from_state(tc, lv, data, self.g, [self.cfg.bs] * self.dim, self.dim,
                                      self.acm.domain_size, self.forest.jmax, self.time, self.iteration)
        paths = []
        for name, f in zip(names, files):
            path = os.path.join(directory, f"{name}_{h5io.timestr(self.time)}.h5")
            h5io.write_wabbit_file(path, f)
            paths.append(path)
        return paths

    def load(self, paths) -> None:
        """Resume from saved files (read_from_files = 1:
        setInitialCondition_tree.f90:52-60 -> readHDF5vct_tree,
        InputOutput.f90:322-764): the files' leaves and data replace the
        current ones, time and iteration are taken from the files.  Ghost
        nodes are filled by the next step's sync_ghosts_tree."""
        from . import h5io
        files = [h5io.read_wabbit_file(p) for p in paths]
        tc, lv, data, time, it = h5io.state_from_files(files, self.g, max_level=self.forest.jmax)
        if data.shape[1] != self.acm.neq:
            raise ValueError(f"{data.shape[1]} component files for {self.acm.neq} equations")
        self.set_blocks(tc, lv, data)
        self.time, self.iteration = float(time), int(it)
        self._dev_time = None

    def mask(self):
        self._fresh()
        tc, lv, hv, _ = self.forest.blocks()
        order = np.lexsort((tc, lv))
        return np.stack([self.pool.download(L.WH_MASK, int(h), 1)[0] for h in hv[order]])

    def close(self):
        self.pool.close()
        self.forest.close()
