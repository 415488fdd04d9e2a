"""GPU: independent adaptive solves sharing pools (wabbit_amd.batch).

A solve's result must not depend on the batch it runs in: the same solve run
alone, run with others (mixed Jmax and Bs, so two pools), and run through the
single-forest driver (AdaptiveACM) gives bit-identical block lists and data.
The per-block path (one launch per block, path 1) matches the fused path
(path 2) to rel. L2 1e-13 (the same arithmetic in the same order)."""
import numpy as np
import pytest

from wabbit_amd import workloads as wl
from wabbit_amd.batch import Batch, run_alone

pytestmark = pytest.mark.gpu

NSTEPS = 6


def specs():
    return [wl.channel_spec(11, 4, 16, NSTEPS, 1024), wl.channel_spec(12, 5, 16, NSTEPS, 2560),
            wl.channel_spec(13, 4, 24, NSTEPS, 1024), wl.channel_spec(14, 5, 24, NSTEPS, 2560)]


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def test_single_solve_matches_adaptive_driver():
    spec = wl.channel_spec(12, 5, 16, NSTEPS, 2560)
    alone = run_alone(spec, 4096)
    w = wl.channel(seed=12, jmax=5, bs=16)
    s = wl.channel_solver(w, nb_max=2560)
    try:
        for _ in range(NSTEPS):
            s.step()
        ref = s.state()
    finally:
        s.close()
    assert _same(alone, ref)


def test_alone_vs_in_mix():
    sp = specs()
    alone = [run_alone(x, 4096) for x in sp]
    print("alone:", [len(a[0]) for a in alone], flush=True)
    b = Batch({(2, 16): 4096, (2, 24): 4096})
    try:
        solves = [b.admit(x) for x in sp]
        assert all(s is not None for s in solves)
        assert len(b.groups) == 2
        while not all(s.done for s in solves):
            b.step()
        mixed = [s.state() for s in solves]
    finally:
        b.close()
    for a, m in zip(alone, mixed):
        assert _same(a, m)


def test_per_block_path_matches_fused():
    sp = specs()[:2]
    out = {}
    for path in (1, 2):
        b = Batch({(2, 16): 4096})
        try:
            solves = [b.admit(x) for x in sp]
            for _ in range(3):
                b.step(path)
            out[path] = [s.state() for s in solves]
        finally:
            b.close()
    for (tc1, lv1, d1), (tc2, lv2, d2) in zip(out[1], out[2]):
        assert np.array_equal(tc1, tc2) and np.array_equal(lv1, lv2)
        assert np.linalg.norm(d1 - d2) <= 1e-13 * np.linalg.norm(d2)


def test_slot_ranges_are_reused():
    """A retired solve's slot range admits the next one (the serving loop's
    admission by free memory)."""
    sp = wl.channel_spec(11, 4, 16, 2, 1024)
    b = Batch({(2, 16): 1024})
    try:
        s1 = b.admit(sp)
        assert s1 is not None and b.admit(sp) is None      # pool full
        while not s1.done:
            b.step()
        res = b.retire_done()
        assert len(res) == 1 and res[0][0] == s1.rid and res[0][3] is None
        s2 = b.admit(sp)
        assert s2 is not None and s2.base == s1.base
        while not s2.done:
            b.step()
        assert _same(res[0][2], s2.state())
    finally:
        b.close()
