"""Time integration over a device block pool, mirroring the reference's call
surface:

  RungeKuttaGeneric(time, dt, iteration, params, hvy_block, hvy_work, ...)
      LIB/TIME/runge_kutta_generic.f90:1-156
  calculate_time_step(params, time, iteration, hvy_block, dt, tree_ID)
      LIB/TIME/calculate_time_step.f90:2-125
  timeStep_tree(...)   LIB/TIME/timeStep_tree.f90:1-66
  main loop            LIB/MAIN/main.f90:293-501

Heavy data never leaves the device; the host keeps time, iteration and the
topology.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import lib as L
from .params import ACMParams, TimeParams, clamp_dt

PATH_REFERENCE = 0   # one launch per operation over all active blocks
PATH_PER_BLOCK = 1   # one launch per block and operation ("plain")
PATH_FUSED = 2       # RHS + RK update fused per stage


@dataclass
class StepInfo:
    time: float
    iteration: int
    dt: float


class AcmSolver:
    """ACM time stepping of one forest on the GPU."""

    def __init__(self, forest, acm: ACMParams, tp: TimeParams, g_rhs: int, path=PATH_FUSED,
                 mask=None, n_mask=0):
        self.forest, self.acm, self.tp, self.g_rhs, self.path = forest, acm, tp, g_rhs, path
        bt = np.asarray(tp.butcher, np.float64)
        slots = 3 if path == PATH_FUSED else bt.shape[0]
        self.pool = L.Pool(forest.dim, forest.bs, forest.g, acm.neq, forest.n, n_work_slots=slots,
                           n_mask=n_mask, mask_per_block=True)
        self.pool.set_active(forest.active, forest.x0dx)
        self.pool.set_patches(0, forest.same_level_patches(g_rhs))
        self.pool.set_patches(1, forest.same_level_patches(forest.g))
        self.acm_c = L.acm_params_c(acm)
        self.butcher = bt
        self.time, self.iteration = 0.0, 0
        if mask is not None:
            self.pool.upload(L.WH_MASK, 0, mask)

    def load(self, hvy, time=0.0, iteration=0):
        self.pool.upload(L.WH_BLOCK, 0, hvy)
        self.time, self.iteration = float(time), int(iteration)

    def state(self):
        return self.pool.download(L.WH_BLOCK, 0, self.forest.n)

    def calculate_time_step(self):
        """dt = clamps(min over blocks of GET_DT_BLOCK_ACM)."""
        return clamp_dt(self.pool.dt(self.acm_c), self.time, self.tp)

    def step(self) -> StepInfo:
        """timeStep_tree: one RK step, then time += dt, iteration += 1."""
        dt = self.calculate_time_step()
        self.pool.rk_step(self.acm_c, self.butcher, self.time, dt, self.path)
        self.time += dt
        self.iteration += 1
        return StepInfo(self.time, self.iteration, dt)

    def run(self, nsteps=None):
        """Main-loop stop rule: time < time_max and iteration < nt
        (main.f90:293)."""
        n = 0
        while self.time < self.tp.time_max and self.iteration < self.tp.nt:
            if nsteps is not None and n >= nsteps:
                break
            self.step()
            n += 1
        return n

    def run_device(self, nsteps, use_graph=True):
        """The same loop with dt, clamps and time kept on the device (and the
        step replayed from a hipGraph)."""
        self.pool.time_set(self.time, self.iteration)
        self.pool.run_steps(self.acm_c, self.butcher, L.time_params_c(self.tp), nsteps, use_graph)
        self.time, self.iteration, _ = self.pool.time_get()
