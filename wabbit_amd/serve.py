"""A request loop over shared device pools: continuous batching of
independent adaptive ACM solves.

Requests arrive over time (a seeded Poisson stream); each is admitted as soon
as its pool has a free range of block slots (the reference's --memory per
solve), its state stays on the device for its whole life, every loop
iteration advances all admitted solves by one main-loop step (their RK
stages in one launch per operation and per pool), and a finished solve is
retired at once, freeing its slots for the next waiting request.  Requests
that do not fit yet wait in arrival order, but a smaller later request may
pass a waiting larger one (no head-of-line blocking).

Because no reduction or launch parameter depends on the batch, a request's
result is the same whether it runs alone or in the mix.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import lib as L
from .batch import Batch, SolveSpec
from .timestep import PATH_FUSED


@dataclass
class Request:
    rid: int
    spec: SolveSpec
    arrival: float                 # seconds after the start of the run
    admitted: Optional[float] = None
    finished: Optional[float] = None
    error: Optional[str] = None
    state: Optional[tuple] = None

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


@dataclass
class ServeStats:
    completed: int
    failed: int
    wall: float
    solves_per_s: float
    p50: float
    p99: float
    mean_wait: float
    block_rk_stages: int
    max_concurrent: int
    per_request: List[dict] = field(default_factory=list)


class Server:
    """Continuous batching over a Batch of shared pools."""

    def __init__(self, slots_per_pool: Dict[tuple, int], n_work_slots: int = 3, path: int = PATH_FUSED,
                 keep_states: bool = False, lanes: int = 1, prealloc=None):
        """prealloc: pool keys (batch.pool_key) created at start-up; then
        admission never creates a pool."""
        self.batch = Batch(slots_per_pool, n_work_slots=n_work_slots, lanes=lanes, prealloc=prealloc)
        self.path = path
        self.keep_states = keep_states

    def run(self, requests: List[Request], clock=time.perf_counter, log=None) -> ServeStats:
        reqs = sorted(requests, key=lambda r: r.arrival)
        by_rid = {r.rid: r for r in reqs}
        pending = list(reqs)            # not yet arrived
        waiting: List[Request] = []     # arrived, not admitted
        t0 = clock()
        stages, max_conc, done = 0, 0, 0
        while pending or waiting or self.batch.solves:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.pop(0))
            for r in list(waiting):
                try:
                    s = self.batch.admit(r.spec, rid=r.rid)
                except L.WabbitHipError as e:   # its initial grid does not fit its slots
                    r.error, r.finished = str(e), clock() - t0
                    waiting.remove(r)
                    if log:
                        log(f"request {r.rid} ({r.spec.tag}) failed at admission: {e}")
                    continue
                if s is not None:
                    r.admitted = clock() - t0
                    waiting.remove(r)
            if not self.batch.solves:
                if pending:   # idle until the next arrival
                    time.sleep(max(0.0, min(0.05, pending[0].arrival - (clock() - t0))))
                    continue
                if waiting:   # nothing running and the head does not fit: it never will
                    for r in waiting:
                        r.error = "does not fit in its pool"
                        r.finished = clock() - t0
                        if log:
                            log(f"request {r.rid} ({r.spec.tag}) does not fit in its pool")
                    waiting = []
                continue
            max_conc = max(max_conc, len(self.batch.solves))
            stages += self.batch.step(self.path)
            for rid, tag, st, err in self.batch.retire_done(keep_state=self.keep_states):
                r = by_rid[rid]
                r.finished = clock() - t0
                r.error = err
                r.state = st
                done += 1
                if log:
                    log(f"request {rid} ({tag}) done at {r.finished:.3f} s, latency {r.latency:.3f} s"
                        + (f", failed: {err}" if err else ""))
        wall = clock() - t0
        ok = [r for r in reqs if r.error is None and r.finished is not None]
        lat = np.array([r.latency for r in ok]) if ok else np.zeros(1)
        waits = np.array([r.admitted - r.arrival for r in ok if r.admitted is not None]) if ok else np.zeros(1)
        return ServeStats(completed=len(ok), failed=len(reqs) - len(ok), wall=wall,
                          solves_per_s=len(ok) / wall if wall > 0 else 0.0,
                          p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)),
                          mean_wait=float(waits.mean()), block_rk_stages=stages, max_concurrent=max_conc,
                          per_request=[dict(rid=r.rid, tag=r.spec.tag, arrival=r.arrival, admitted=r.admitted,
                                            finished=r.finished, error=r.error) for r in reqs])

    def close(self):
        self.batch.close()
