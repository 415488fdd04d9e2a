"""Many independent adaptive ACM solves on shared device pools.

The reference runs one forest per process (LIB/MAIN/main.f90); here a batch of
independent solves -- an ensemble, or the requests of a serving loop -- share
one block pool per block shape (dim, Bs, wavelet, scheme, masked) and host
thread ("lane").  Per main-loop iteration:

  A. per solve, on its lane's pool (lanes in parallel): sync_ghosts_tree,
     refine_tree "everywhere", createMask_tree, the RHS ghost patches; per
     pool the union active list with a group index per block
     (wh_set_block_groups), each group's physics constants
     (wh_set_group_params) and the block-dt minimum per group
     (wh_dt_reduce_groups), clamped per solve on the host;
  B. ONE RungeKuttaGeneric step for all pools of a dimension
     (wh_batch_rk_step): per RK stage one ghost fill per sync stage and one
     fused stage kernel over every block of every pool, replayed from a
     hipGraph;
  C. per solve (lanes in parallel): the time advance and adapt_tree.

Mesh work runs per solve, each inside its own range of pool slots
(wh_forest_set_slot_base).  Each solve's arithmetic is exactly that of
AdaptiveACM on its own, so a solve's result does not depend on what else is
in the batch.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from . import lib as L
from . import wavelets as W
from .adaptive import AdaptConfig, fd_reach
from .forest import equidistant_treecodes, refined_treecodes
from .params import ACMParams, TimeParams, butcher_rk4, clamp_dt
from .timestep import PATH_FUSED
from .tree import Forest, refine_on_device


@dataclass
class SolveSpec:
    """One independent adaptive ACM solve."""
    acm: ACMParams
    cfg: AdaptConfig
    inicond: Callable            # (treecodes, levels, jmax_treecode) -> (n, neq, nz, ny, nx)
    nsteps: int
    mask: Optional[dict] = None  # wavelets.mask_params keywords (penalised solves)
    tp: TimeParams = field(default_factory=lambda: TimeParams(time_max=1e9))
    slots: int = 4096            # pool slots reserved for the solve (its --memory)
    jini: Optional[int] = None   # initial equidistant level (default Jmin)
    tag: str = ""


class SlotRanges:
    """First-fit allocator of contiguous slot ranges of one pool."""

    def __init__(self, n):
        self.n = n
        self.used: Dict[int, int] = {}   # base -> count

    def alloc(self, count):
        pos = 0
        for b in sorted(self.used):
            if b - pos >= count:
                break
            pos = b + self.used[b]
        if pos + count > self.n:
            return None
        self.used[pos] = count
        return pos

    def free(self, base):
        del self.used[base]

    @property
    def in_use(self):
        return sum(self.used.values())


def pool_key(spec: "SolveSpec"):
    """What a pool is shared by: the block shape and the kernels' template
    choices; physics constants travel per group."""
    return (spec.acm.dim, spec.cfg.bs, spec.cfg.wavelet, spec.mask is not None, spec.acm.scheme,
            bool(spec.acm.skew_symmetry))


def slot_bytes(key, n_work_slots: int = 3) -> int:
    """Device bytes one pool slot of pool_key `key` holds: hvy_block, the
    work slots and hvy_tmp (neq components each) plus the 6-component mask
    of penalised pools."""
    dim, bs, wavelet, masked, scheme, _ = key
    g = W.setup(wavelet, fd_reach(scheme)).g
    cs = (bs + 2 * g) ** dim
    neq = dim + 1
    return 8 * cs * (neq * (2 + n_work_slots) + (6 if masked else 0))


def budget_slots(keys, weights, budget_bytes: float, lanes: int = 1, n_work_slots: int = 3) -> Dict[tuple, int]:
    """Slots per (dim, Bs) for pools created at start-up within budget_bytes:
    key k gets weights[k] of the budget, rounded down to a multiple of
    lanes (every lane's pool set holds 1/lanes of each pool)."""
    out = {}
    tot = sum(weights[k] for k in keys)
    for k in keys:
        n = int(budget_bytes * weights[k] / tot // slot_bytes(k, n_work_slots))
        out[(k[0], k[1])] = n - n % lanes
    return out


class Solve:
    """A solve admitted to a pool group."""

    def __init__(self, rid, spec: SolveSpec, group: "PoolGroup", base: int):
        self.rid, self.spec, self.group, self.base = rid, spec, group, base
        self.time, self.iteration = 0.0, 0
        self.rk_blocks = 0
        self.error: Optional[str] = None   # set when the solve failed (e.g. out of its slots)
        cfg, acm = spec.cfg, spec.acm
        self.acm_c = L.acm_params_c(acm)
        self.forest = Forest(acm.dim, cfg.jmax, [cfg.bs] * acm.dim, group.w.g, acm.domain_size[:acm.dim],
                             spec.slots, slot_base=base)
        self.mask_p = W.mask_params(**spec.mask) if spec.mask is not None else None
        lifted = bool(group.w.lifted)
        self.adapt_p = W.adapt_params(
            cfg.jmin, cfg.jmax, W.WH_IND_THRESHOLD,
            use_ce=lifted if cfg.use_ce is None else cfg.use_ce, eps=cfg.eps, groups=cfg.groups,
            force_maxlevel_dealiasing=cfg.force_maxlevel_dealiasing, eps_normalized=cfg.eps_normalized,
            threshold_mask=cfg.threshold_mask and self.mask_p is not None,
            security_zone=lifted if cfg.security_zone is None else cfg.security_zone,
            domain=acm.domain_size, mask=self.mask_p)

    @property
    def done(self):
        return self.error is not None or self.iteration >= self.spec.nsteps

    def _set_blocks(self, tc, lv, data):
        n = len(tc)
        if n > self.spec.slots:
            raise L.WabbitHipError(-5, f"solve {self.rid}: {n} blocks exceed its {self.spec.slots} slots")
        self.forest.set_blocks(tc, lv, self.base + np.arange(n, dtype=np.int32))
        self.group.pool.upload(L.WH_BLOCK, self.base, data)

    def initial_grid(self):
        """setInitialCondition_tree.f90:75-108 for an analytic initial condition."""
        cfg, dim, ic = self.spec.cfg, self.spec.acm.dim, self.spec.inicond
        jtc = self.forest.jmax
        tc, lv = equidistant_treecodes(cfg.jmin if self.spec.jini is None else self.spec.jini, dim, jtc)
        self._set_blocks(tc, lv, ic(tc, lv, jtc))
        n_old, it = -1, 0
        while self.forest.n != n_old and it < cfg.jmax - cfg.jmin:
            n_old = self.forest.n
            tc, lv, _, _ = self.forest.blocks()
            tc, lv = refined_treecodes(tc, lv, jtc, dim, cfg.jmax)
            self._set_blocks(tc, lv, ic(tc, lv, jtc))
            W.adapt_tree(self.group.pool, self.forest, self.adapt_p)
            it += 1

    def state(self):
        """(treecodes, levels, data) of the leaves, sorted by (level, treecode)."""
        tc, lv, hv, _ = self.forest.blocks()
        order = np.lexsort((tc, lv))
        pool = self.group.pool
        data = np.stack([pool.download(L.WH_BLOCK, int(h), 1)[0] for h in hv[order]])
        return tc[order], lv[order], data

    def close(self):
        self.forest.close()


class PoolGroup:
    """One shared pool for the solves of one block shape (pool_key)."""

    def __init__(self, key, n_slots: int, n_work_slots: int = 3):
        """n_work_slots: 3 suffice for the fused path and adapt_tree, the
        per-block path needs 5 (RK4)."""
        dim, bs, wavelet, masked, scheme, skew = key
        self.key = key
        self.dim = dim
        self.w = W.setup(wavelet, fd_reach(scheme))
        neq = dim + 1
        self.pool = L.Pool(dim, [bs] * dim, self.w.g, neq, n_slots, n_work_slots=n_work_slots,
                           n_tmp_eqn=neq, n_mask=6 if masked else 0, mask_per_block=True)
        W.set_wavelet(self.pool, self.w)
        self.slots = SlotRanges(n_slots)
        self.solves: List[Solve] = []
        self.last_stages = 0
        self._live: List[Solve] = []
        self._dts = np.zeros(0)

    def admit(self, rid, spec: SolveSpec) -> Optional[Solve]:
        base = self.slots.alloc(spec.slots)
        if base is None:
            return None
        s = Solve(rid, spec, self, base)
        try:
            s.initial_grid()
        except Exception:
            s.close()
            self.slots.free(base)
            raise
        self.solves.append(s)
        return s

    def retire(self, s: Solve):
        self.solves.remove(s)
        s.close()
        self.slots.free(s.base)

    def prepare(self):
        """Phase A: sync_ghosts_tree + refine everywhere + mask of every
        solve (each a call for all solves of the pool), then the pool's union
        active list, patches, groups, group constants and clamped dt per
        group.  A pool without live solves gets empty tables."""
        pool, w = self.pool, self.w
        cand = [s for s in self.solves if not s.done]
        # sync_ghosts_tree of every solve, one launch per operation for all of
        # them; should that fail, each solve alone, so a failure stays its own
        try:
            W.sync_trees(pool, [s.forest for s in cand])
        except L.WabbitHipError:
            for s in cand:
                try:
                    W.sync_tree(pool, s.forest)
                except L.WabbitHipError as e:
                    s.error = str(e)
            cand = [s for s in cand if s.error is None]
        # refine everywhere below Jmax: every solve's new blocks on the host
        # threads and one refinement launch for all, in one call; a solve
        # whose free slots cannot take all its daughters at once refines in
        # batches of its own
        def fail(s, code, msg):
            s.error = f"{L.ERRORS.get(code, code)}: {msg}"
        res = W.refine_trees(pool, [s.forest for s in cand], [s.spec.cfg.jmax for s in cand], w.order)
        for s, (code, msg) in zip(list(cand), res):
            if code == L.WH_E_NOMEM:
                try:
                    tc, lv, hv, _ = s.forest.blocks()
                    refine_on_device(pool, s.forest, np.asarray(hv[lv < s.spec.cfg.jmax], np.int32), w.order,
                                     w.g_rhs, presync=None, attach=False)
                    continue
                except L.WabbitHipError as e:   # the solve alone fails
                    code, msg = e.code, str(e)
            if code:
                fail(s, code, msg)
                cand.remove(s)
        masked = [s for s in cand if s.mask_p is not None]
        for s, (code, msg) in zip(masked, W.create_masks(pool, [s.forest for s in masked],
                                                         [s.mask_p for s in masked])):
            if code:
                fail(s, code, msg)
                cand.remove(s)
        # the RK step's tables from the forests (active list, origins and
        # spacings, one block group per solve, RHS ghost patches)
        res, nbs = W.set_tables_from_forests(pool, [s.forest for s in cand], w.g_rhs, w.order)
        live = []
        for s, (code, msg), nb in zip(cand, res, nbs):
            if code:
                fail(s, code, msg)
                continue
            s.rk_blocks = nb
            live.append(s)
        self._live = live
        self.last_nact = int(sum(nbs))
        if not live:
            self._dts = np.zeros(0)
            return
        pool.set_group_params([s.acm_c for s in live])
        dts = pool.dt_groups(live[0].acm_c)
        self._dts = np.array([clamp_dt(d, s.time, s.spec.tp) for d, s in zip(dts, live)])

    def finish(self):
        """Phase C: time advance and adapt_tree per solve; returns the
        block-RK-stages of the step."""
        stages = 0
        for s, dt in zip(self._live, self._dts):
            s.time += float(dt)
            s.iteration += 1
            stages += 4 * s.rk_blocks
        self.last_stages = stages
        # adapt_tree of every solve, their device operations launched together
        live = [s for s in self._live if s.error is None]
        try:
            res = W.adapt_trees(self.pool, [s.forest for s in live], [s.adapt_p for s in live])
        except L.WabbitHipError as e:   # an operation shared by all of them failed
            res = [(e.code, str(e))] * len(live)
        for s, (code, msg) in zip(live, res):
            if code:
                s.error = f"{L.ERRORS.get(code, code)}: {msg}"
        self._live = []
        return stages

    def step_alone(self, path=PATH_FUSED, butcher=None):
        """One main-loop iteration of the group's solves with the pool's own
        RK launches (path 1: one launch per block; path 2: the fused stages
        of this pool only)."""
        self.prepare()
        if self._live:
            bt = butcher_rk4() if butcher is None else butcher
            if any(_differs(s.acm_c, self._live[0].acm_c) for s in self._live):
                raise ValueError("solves with different constants in one pool need the batch step")
            self.pool.set_group_params([])
            self.pool.rk_step_groups(self._live[0].acm_c, bt, self._dts, path)
        return self.finish()

    def close(self):
        for s in list(self.solves):
            s.close()
        self.solves = []
        self.pool.close()


def _differs(a, b):
    return bytes(a) != bytes(b)


class Batch:
    """Independent solves on shared pools, one pool per block shape and lane.

    With lanes > 1 the solves are spread over that many independent sets of
    pools.  Every pool has its own stream, and the mesh work of its solves
    runs on a host thread of its own (in the library, with the interpreter
    lock released); the RK stages of all pools of a dimension run as one
    launch per stage (wh_batch_rk_step) between the two mesh phases.

    prealloc: the pool keys (pool_key) to create at start-up in every lane;
    then admission never creates a pool, and a request whose shape has no
    pool is refused."""

    def __init__(self, slots_per_pool: Dict[tuple, int], n_work_slots: int = 5, lanes: int = 1,
                 prealloc: Optional[List[tuple]] = None, use_graph: bool = True, parallel: bool = True):
        """slots_per_pool: {(dim, Bs): pool slots} -- the device memory set
        aside per shape (the reference's --memory), split evenly over the
        lanes.  parallel=False steps the pools one after the other on the
        calling thread (the same device work; for tracing tools that do not
        follow several host threads calling the runtime)."""
        self.slots_per_pool = dict(slots_per_pool)
        self.n_work_slots = n_work_slots
        self.lanes = max(1, int(lanes))
        self.use_graph = use_graph
        self.parallel = parallel
        self.groups: Dict[tuple, PoolGroup] = {}
        self.next_rid = 0
        self._exec = None
        self._exec_n = 0
        self._steppers: Dict[int, L.BatchStepper] = {}
        self._stepper_pools: Dict[int, tuple] = {}
        self.phase_s: Dict[str, float] = {}
        self.preallocated = prealloc is not None
        for k in prealloc or []:
            for lane in range(self.lanes):
                self._create(k, lane)

    def _create(self, pkey, lane):
        n = self.slots_per_pool[(pkey[0], pkey[1])] // self.lanes
        self.groups[(lane,) + tuple(pkey)] = PoolGroup(pkey, n, self.n_work_slots)

    def key(self, spec: SolveSpec, lane: int = 0):
        return (lane,) + pool_key(spec)

    def group_for(self, spec: SolveSpec, lane: int = 0) -> PoolGroup:
        k = self.key(spec, lane)
        if k not in self.groups:
            if self.preallocated:
                raise ValueError(f"no pool for {pool_key(spec)}: the batch was preallocated without it")
            self._create(pool_key(spec), lane)
        return self.groups[k]

    def admit(self, spec: SolveSpec, rid=None) -> Optional[Solve]:
        """Admit a solve into the least loaded lane whose pool has a free
        slot range; None if none has."""
        if rid is None:
            rid = self.next_rid
            self.next_rid += 1

        def load(lane):
            g = self.groups.get(self.key(spec, lane))
            return g.slots.in_use if g is not None else 0
        for lane in sorted(range(self.lanes), key=load):
            s = self.group_for(spec, lane).admit(rid, spec)
            if s is not None:
                return s
        return None

    @property
    def solves(self):
        return [s for g in self.groups.values() for s in g.solves]

    def _run_groups(self, fn):
        """fn(group) for every pool group, each on a host thread of its own
        (the groups share nothing but the device); the results in group
        order."""
        gs = list(self.groups.values())
        if len(gs) <= 1 or not self.parallel:
            return [fn(g) for g in gs]
        if self._exec_n < len(gs):
            from concurrent.futures import ThreadPoolExecutor
            if self._exec is not None:
                self._exec.shutdown()
            self._exec = ThreadPoolExecutor(len(gs))
            self._exec_n = len(gs)
        futs = [self._exec.submit(fn, g) for g in gs]
        return [f.result() for f in futs]

    def _stepper(self, dim, groups):
        pools = tuple(g.pool for g in groups)
        ids = tuple(id(p) for p in pools)
        if self._stepper_pools.get(dim) != ids:
            if dim in self._steppers:
                self._steppers[dim].close()
            self._steppers[dim] = L.BatchStepper(pools)
            self._stepper_pools[dim] = ids
        return self._steppers[dim]

    def step(self, path=PATH_FUSED):
        """One iteration of every admitted solve; returns block-RK-stages.
        Wall time of the three phases accumulates in phase_s (prepare: the
        mesh work before the RK step, rk: the batched RK launches, finish:
        adapt_tree)."""
        if path != PATH_FUSED:
            return sum(self._run_groups(lambda g: g.step_alone(path)))
        t0 = time.perf_counter()
        self._run_groups(lambda g: g.prepare())
        t1 = time.perf_counter()
        bt = butcher_rk4()
        for dim in (2, 3):
            groups = [g for g in self.groups.values() if g.dim == dim]
            if not any(g._live for g in groups):
                continue
            live = [g for g in groups if g._live]
            st = self._stepper(dim, groups)
            dts = np.concatenate([g._dts for g in groups]) if groups else np.zeros(0)
            st.rk_step(live[0]._live[0].acm_c, bt, dts, use_graph=self.use_graph)
        t2 = time.perf_counter()
        stages = sum(self._run_groups(lambda g: g.finish()))
        t3 = time.perf_counter()
        for k, v in (("prepare", t1 - t0), ("rk", t2 - t1), ("finish", t3 - t2)):
            self.phase_s[k] = self.phase_s.get(k, 0.0) + v
        return stages

    def stage_kernel_roofline(self, reps=5):
        """{dim: (ms per launch, blocks per launch)} of the batched fused
        stage kernel on the last step's union grid."""
        out = {}
        bt = butcher_rk4()
        for dim, st in self._steppers.items():
            acm = next((s.acm_c for g in self.groups.values() if g.dim == dim for s in g.solves), None)
            if acm is not None:
                out[dim] = st.time_stage_kernel(acm, bt, reps)
        return out

    def retire_done(self, keep_state=True):
        """Remove finished (or failed) solves; returns [(rid, tag, state or
        None, error or None)]."""
        out = []
        for g in self.groups.values():
            for s in [s for s in g.solves if s.done]:
                st = s.state() if keep_state and s.error is None else None
                out.append((s.rid, s.spec.tag, st, s.error))
                g.retire(s)
        return out

    def close(self):
        for st in self._steppers.values():
            st.close()
        self._steppers = {}
        self._stepper_pools = {}
        for g in self.groups.values():
            g.close()
        self.groups = {}
        if self._exec is not None:
            self._exec.shutdown()
            self._exec = None


def run_alone(spec: SolveSpec, slots: int, path=PATH_FUSED):
    """The solve on a batch of its own (the reference's one-forest run)."""
    b = Batch({(spec.acm.dim, spec.cfg.bs): slots})
    try:
        s = b.admit(spec)
        while not s.done:
            b.step(path)
        return s.state()
    finally:
        b.close()

