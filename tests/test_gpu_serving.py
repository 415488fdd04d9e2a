"""GPU: the request loop (wabbit_amd.serve) with continuous batching.

A small stream of requests with staggered arrivals -- 2D penalised channel
solves of two block sizes and a 3D periodic solve -- on shared pools split
over two host lanes.  Requests arriving while others run are admitted as
slots allow; each result is bit-identical to the same solve run alone (the
bar: equal block lists and equal data)."""
import numpy as np
import pytest

from wabbit_amd import workloads as wl
from wabbit_amd.batch import run_alone
from wabbit_amd.serve import Request, Server

pytestmark = pytest.mark.gpu


def stream():
    return [Request(0, wl.channel_spec(21, 4, 16, 5, 1024), 0.0),
            Request(1, wl.channel_spec(22, 5, 24, 4, 2560), 0.0),
            Request(2, wl.periodic3d_spec(23, 3, 16, 2, 768), 0.05),
            Request(3, wl.channel_spec(24, 5, 16, 3, 2560), 0.1),
            Request(4, wl.channel_spec(25, 4, 16, 6, 1024), 0.15)]


def test_results_alone_and_in_mix():
    reqs = stream()
    srv = Server({(2, 16): 8192, (2, 24): 8192, (3, 16): 1536}, n_work_slots=3, keep_states=True, lanes=2)
    try:
        st = srv.run(reqs, log=lambda m: print(m, flush=True))
    finally:
        srv.close()
    assert st.completed == len(reqs) and st.failed == 0
    assert st.max_concurrent >= 2
    assert 0.0 <= st.p50 <= st.p99
    for r in reqs:
        alone = run_alone(r.spec, r.spec.slots)
        assert all(np.array_equal(a, b) for a, b in zip(alone, r.state)), r.spec.tag


def test_request_that_cannot_fit_fails_alone():
    """A request whose slots exceed its pool is reported as failed; the
    others complete."""
    reqs = [Request(0, wl.channel_spec(31, 4, 16, 2, 1024), 0.0),
            Request(1, wl.channel_spec(32, 4, 16, 2, 4096), 0.0)]
    srv = Server({(2, 16): 2048}, n_work_slots=3)
    try:
        st = srv.run(reqs)
    finally:
        srv.close()
    assert st.completed == 1 and st.failed == 1
    assert reqs[1].error is not None and reqs[0].error is None
