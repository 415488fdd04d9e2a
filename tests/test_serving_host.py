"""Host logic of the serving workload (no GPU): the request stream."""
import numpy as np

from wabbit_amd import workloads as wl


def test_serving_mix_shape():
    mix = wl.serving_mix(n=256, seed=2024, rate=4.0)
    assert len(mix) == 256
    t = np.array([a for _, _, a in mix])
    assert np.all(np.diff(t) > 0)
    dims = np.array([s.acm.dim for _, s, _ in mix])
    assert 0.6 < np.mean(dims == 2) < 0.8
    for _, s, _ in mix:
        if s.acm.dim == 2:
            assert s.cfg.jmax in (4, 5, 6, 7) and s.cfg.bs in (16, 24) and 100 <= s.nsteps <= 500
            assert s.mask is not None
        else:
            assert s.cfg.jmax in (4, 5) and s.cfg.bs == 24 and 50 <= s.nsteps <= 200
            assert s.mask is None
    # mean inter-arrival time close to 1/rate
    assert abs(np.mean(np.diff(t)) - 0.25) < 0.05
