"""Host logic of the batched path: the slot-range allocator and the workload
specs (no GPU)."""
from wabbit_amd import workloads as wl
from wabbit_amd.batch import Batch, SlotRanges


def test_slot_ranges_first_fit():
    r = SlotRanges(100)
    a = r.alloc(40)
    b = r.alloc(40)
    assert (a, b) == (0, 40) and r.alloc(30) is None
    r.free(a)
    assert r.alloc(50) is None and r.alloc(30) == 0 and r.alloc(10) == 30
    assert r.alloc(20) == 80 and r.in_use == 100
    r.free(40)
    assert r.alloc(40) == 40


def test_pool_keys_split_shapes_and_physics():
    b = Batch({(2, 16): 10, (2, 24): 10})
    s16 = wl.channel_spec(1, 4, 16, 1, 8)
    s24 = wl.channel_spec(2, 4, 24, 1, 8)
    s16b = wl.channel_spec(3, 5, 16, 1, 8)
    assert b.key(s16) != b.key(s24)
    assert b.key(s16) == b.key(s16b)          # same shape, different geometry and Jmax
    s16c = wl.channel_spec(4, 4, 16, 1, 8)
    s16c.acm.c_0, s16c.acm.nu = 10.0, 2e-3
    assert b.key(s16) == b.key(s16c)          # physics constants travel per group, not per pool


def test_ensemble_members():
    sp = wl.ensemble(64)
    assert len(sp) == 64
    assert {s.cfg.jmax for s in sp} == {4, 5, 6, 7} and {s.cfg.bs for s in sp} == {16, 24}
    assert len({tuple(s.mask["x_cntr"]) for s in sp}) == 64
    assert all(0.1 <= s.mask["R_cyl"] <= 0.2 for s in sp)
