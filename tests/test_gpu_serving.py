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


def test_stream_with_admission_waits_matches_oracle():
    """Pools preallocated at start-up and too small for all requests at once:
    requests wait for slots (continuous batching admits them as earlier ones
    retire).  Every result equals its run alone, bit for bit, and two of
    them (a 2D channel solve and a 3D periodic solve) equal the oracle's
    main loop from the same initial grid (block lists identical, fields
    within relative L2 1e-12)."""
    import periodic3d_case as PC
    from test_gpu_batch_oracle import _oracle_run

    from wabbit_amd.batch import pool_key
    specs = [wl.channel_spec(71, 5, 16, 4, 2560), wl.channel_spec(72, 4, 16, 3, 1024),
             wl.channel_spec(73, 4, 24, 3, 1024), wl.channel_spec(74, 4, 24, 2, 1024),
             wl.periodic3d_spec(75, 3, 16, 2, 768), wl.periodic3d_spec(76, 3, 16, 2, 768),
             wl.channel_spec(77, 4, 16, 2, 1024)]
    reqs = [Request(i, sp, 0.02 * i) for i, sp in enumerate(specs)]
    # one J5 or two J4 channel solves per 2D pool, one 3D solve at a time
    slots = {(2, 16): 2560, (2, 24): 1024, (3, 16): 768}
    srv = Server(slots, n_work_slots=3, keep_states=True, lanes=1,
                 prealloc=sorted({pool_key(sp) for sp in specs}))
    try:
        st = srv.run(reqs, log=lambda m: print(m, flush=True))
    finally:
        srv.close()
    assert st.completed == len(reqs) and st.failed == 0
    waited = [r for r in reqs if r.admitted - r.arrival > 0.05]
    print(f"{len(waited)} of {len(reqs)} requests waited for slots; p50 {st.p50:.3f} s p99 {st.p99:.3f} s",
          flush=True)
    assert len(waited) >= 2
    for r in reqs:
        alone = run_alone(r.spec, r.spec.slots)
        assert all(np.array_equal(a, b) for a, b in zip(alone, r.state)), r.spec.tag
    # the oracle on a 2D and a 3D request, from the initial grids the solves start from
    from wabbit_amd.batch import Batch
    for r in (reqs[1], reqs[4]):
        b = Batch({(r.spec.acm.dim, r.spec.cfg.bs): r.spec.slots}, n_work_slots=3)
        try:
            s = b.admit(r.spec)
            st0 = s.state()
        finally:
            b.close()
        tc, lv, data = r.state
        if r.spec.acm.dim == 2:
            (tco, lvo, want), _ = _oracle_run(r.spec, st0, r.spec.nsteps)
        else:
            w = wl.periodic3d(seed=75, jmax=3, bs=16)
            tree, _ = PC.oracle_run(w, *st0, r.spec.nsteps)
            tco, lvo, want = tree.get()
            tree.close()
        assert np.array_equal(tc, tco) and np.array_equal(lv, lvo), r.spec.tag
        g, bs, dim = (data.shape[-1] - r.spec.cfg.bs) // 2, r.spec.cfg.bs, r.spec.acm.dim
        inner = (slice(None), slice(None)) + ((0,) if dim == 2 else (slice(g, g + bs),)) + \
            (slice(g, g + bs), slice(g, g + bs))
        err = np.linalg.norm(data[inner] - want[inner]) / np.linalg.norm(want[inner])
        print(r.spec.tag, len(tc), "blocks, rel L2 vs oracle", err, flush=True)
        assert err <= 1e-12, r.spec.tag
