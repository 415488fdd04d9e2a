"""Host forest over the C ABI (wh_forest_*): the block list of one tree and
the device tables built from it.

This is the stand-in for the reference's host metadata (lgt_block,
hvy_neighbor, hvy_active; LIB/MESH/module_forestMetaData.f90:33-56) on the
path: block list -> ghost-patch tables (sync_ghosts_RHS_tree /
sync_ghosts_tree on leaf grids), origins/spacings, refinement bookkeeping.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib as L

_vp = ctypes.c_void_p
_BOUND = False


def _bind():
    global _BOUND
    if _BOUND:
        return
    lib = L.load()
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "wh_forest_create": [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int, _vp,
                             ctypes.c_int, _vp, _vp, ctypes.c_int],
        "wh_forest_destroy": [_vp],
        "wh_forest_set_slot_base": [_vp, ctypes.c_int],
        "wh_forest_set_blocks": [_vp, ctypes.c_int, _vp, _vp, _vp],
        "wh_forest_size": [_vp, ip],
        "wh_forest_get_blocks": [_vp, ctypes.c_int, _vp, _vp, _vp, _vp],
        "wh_forest_patches": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, ip],
        "wh_forest_refine": [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, ip],
        "wh_refine_blocks": [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _BOUND = True


class Forest:
    """One tree's leaf blocks (treecode, level, hvy id) on a pool of nb_max slots."""

    def __init__(self, dim, jmax, bs, g, domain, nb_max, periodic=(1, 1, 1), slot_base=0):
        """slot_base: the forest uses pool slots [slot_base, slot_base + nb_max)."""
        _bind()
        self.dim, self.jmax, self.g, self.nb_max = dim, jmax, g, nb_max
        self.slot_base = slot_base
        self.bs = list(bs) + [1] * (3 - len(bs))
        self.domain = list(domain) + [1.0] * (3 - len(domain))
        b = np.asarray(self.bs, np.int32)
        d = np.asarray(self.domain, np.float64)
        p = np.asarray(list(periodic) + [1] * (3 - len(periodic)), np.int32)
        h = ctypes.c_void_p()
        L.check(L.load().wh_forest_create(ctypes.byref(h), dim, jmax, b.ctypes.data, g, d.ctypes.data,
                                          p.ctypes.data, nb_max))
        self.h = h
        if slot_base:
            L.check(L.load().wh_forest_set_slot_base(h, slot_base))

    def close(self):
        if getattr(self, "h", None):
            L.load().wh_forest_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_blocks(self, treecodes, levels, hvy_ids):
        tc = np.ascontiguousarray(treecodes, np.int64)
        lv = np.ascontiguousarray(levels, np.int32)
        hv = np.ascontiguousarray(hvy_ids, np.int32)
        L.check(L.load().wh_forest_set_blocks(self.h, tc.shape[0], tc.ctypes.data, lv.ctypes.data,
                                              hv.ctypes.data))

    @property
    def n(self):
        n = ctypes.c_int()
        L.check(L.load().wh_forest_size(self.h, ctypes.byref(n)))
        return n.value

    def blocks(self):
        """(treecodes, levels, hvy ids, x0dx (n,6)) in list order."""
        n = self.n
        tc = np.zeros(n, np.int64)
        lv = np.zeros(n, np.int32)
        hv = np.zeros(n, np.int32)
        x = np.zeros((n, 6), np.float64)
        L.check(L.load().wh_forest_get_blocks(self.h, n, tc.ctypes.data, lv.ctypes.data, hv.ctypes.data,
                                              x.ctypes.data))
        return tc, lv, hv, x

    def patches(self, gminus, gplus, order):
        n = ctypes.c_int()
        L.check(L.load().wh_forest_patches(self.h, gminus, gplus, order, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, dtype=L.PATCH_DTYPE)
        L.check(L.load().wh_forest_patches(self.h, gminus, gplus, order,
                                           out.ctypes.data if n.value else None, n.value, ctypes.byref(n)))
        return out

    def refine(self, hvy_ids):
        """Refine blocks (by hvy id); returns (mother, daughter, digit) triples."""
        ids = np.ascontiguousarray(hvy_ids, np.int32)
        cap = ids.shape[0] * (1 << self.dim)
        tri = np.zeros(3 * max(cap, 1), np.int32)
        n = ctypes.c_int()
        L.check(L.load().wh_forest_refine(self.h, ids.shape[0], ids.ctypes.data, tri.ctypes.data, cap,
                                          ctypes.byref(n)))
        return tri[:3 * n.value].reshape(-1, 3)

    def attach(self, pool, g_rhs, order, full=True):
        """Upload the active list, origins/spacings and the patch tables
        (case 0: RHS depth g_rhs; case 1: full depth g, unless full=False)
        to a device pool."""
        tc, lv, hv, x = self.blocks()
        pool.set_active(hv, x)
        pool.set_patches(0, self.patches(g_rhs, g_rhs, order))
        if full:
            pool.set_patches(1, self.patches(self.g, self.g, order))
        return tc, lv, hv, x


def refine_blocks(pool, triples, order, array=L.WH_BLOCK):
    """refineBlock on the device for (mother, daughter, digit) triples of
    Forest.refine (of one or several forests sharing the pool)."""
    _bind()
    t = np.ascontiguousarray(triples, np.int32).reshape(-1, 3)
    if t.shape[0]:
        L.check(L.load().wh_refine_blocks(pool.h, array, t.ctypes.data, t.shape[0], order))


def refine_on_device(pool, forest, hvy_ids, order, g_rhs, array=L.WH_BLOCK, presync="table", attach=True,
                     full=True):
    """refine_tree's execution step for the given blocks: sync the full ghost
    layer (sync_ghosts_tree), refine on the device, then update the pool's
    tables.  presync "table" runs the pool's full-depth patch table (plain
    restriction), "tree" the forest sync with HD-filtered restriction of
    lifted wavelets (wh_sync_tree), None skips the sync.  attach=False leaves
    the pool's tables alone (forests sharing a pool)."""
    if presync == "table":
        pool.sync(array, 1)
    elif presync == "tree":
        from .wavelets import sync_tree
        sync_tree(pool, forest, array)
    ids = [int(h) for h in hvy_ids]
    nd = 1 << forest.dim
    done = []
    # daughters go to free slots (a kernel cannot overwrite a mother it still
    # reads); refine in batches that fit, the mothers' slots are freed after
    # each batch and reused by the next -- the pool needs (2^dim - 1) slots
    # per refined block plus one batch of 2^dim, as the reference's
    # number_blocks sizing for refinement
    while ids:
        free = forest.nb_max - forest.n
        m = min(len(ids), free // nd)
        if m < 1:
            raise L.WabbitHipError(-5, "pool out of memory for refinement")
        tri = forest.refine(np.asarray(ids[:m], np.int32))
        t = np.ascontiguousarray(tri, np.int32)
        L.check(L.load().wh_refine_blocks(pool.h, array, t.ctypes.data, t.shape[0], order))
        done.append(tri)
        ids = ids[m:]
    if attach:
        forest.attach(pool, g_rhs, order, full)
    return np.concatenate(done) if done else np.zeros((0, 3), np.int32)
