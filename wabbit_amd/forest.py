"""Host-side forest metadata for the device block pool.

The reference keeps the tree on the host (lgt_block, treecodes, hvy_neighbor,
hvy_active); this module holds what the device path needs from it: the
active list, per-block origin/spacing and the ghost-patch tables.

Treecodes follow LIB/TREE/module_treelib.f90: an int64 with ``dim`` bits per
level, the digit of level l at bit offset (Jmax-l)*dim, digit bits
(x, y, z) = (bit 1, bit 0, bit 2) -- refinementExecute.f90:98-105.
"""
from __future__ import annotations

import itertools

import numpy as np

from .lib import PATCH_DTYPE, WH_PATCH_COPY


def treecode_from_ijk(ijk, level, jmax, dim):
    """Numerical treecode of the block with 0-based grid index ijk at ``level``
    (encoding_b in module_treelib.f90)."""
    tc = 0
    for lev in range(1, level + 1):
        sh = level - lev
        bx = (ijk[0] >> sh) & 1
        by = (ijk[1] >> sh) & 1
        bz = (ijk[2] >> sh) & 1 if dim == 3 else 0
        digit = by + 2 * bx + 4 * bz
        tc |= digit << ((jmax - lev) * dim)
    return tc


def ijk_from_treecode(tc, level, jmax, dim):
    ijk = [0, 0, 0]
    for lev in range(1, level + 1):
        digit = (tc >> ((jmax - lev) * dim)) & ((1 << dim) - 1)
        ijk[0] = 2 * ijk[0] + ((digit >> 1) & 1)
        ijk[1] = 2 * ijk[1] + (digit & 1)
        ijk[2] = 2 * ijk[2] + ((digit >> 2) & 1 if dim == 3 else 0)
    return ijk


def spacing_origin(ijk, level, domain, bs, dim):
    """get_block_spacing_origin_b (module_treelib.f90:53-103)."""
    dx = [0.0, 0.0, 0.0]
    x0 = [0.0, 0.0, 0.0]
    for d in range(dim):
        dx[d] = 2.0 ** (-level) * domain[d] / float(bs[d])
        x0[d] = float(ijk[d] * bs[d]) * dx[d]
    return x0, dx


class EquidistantForest:
    """All blocks at one level J on a periodic domain (createEquidistantGrid_tree).

    Block k (0-based hvy id) has grid index ijk = unravel(k) with x fastest."""

    def __init__(self, dim, level, bs, g, domain, jmax=None):
        self.dim, self.level, self.g = dim, level, g
        self.bs = list(bs) + [1] * (3 - len(bs))
        self.domain = list(domain) + [1.0] * (3 - len(domain))
        self.jmax = level if jmax is None else jmax
        self.nb = 2 ** level
        nbz = self.nb if dim == 3 else 1
        self.ijk = [(i, j, k) for k in range(nbz) for j in range(self.nb) for i in range(self.nb)]
        self.n = len(self.ijk)
        self.active = np.arange(self.n, dtype=np.int32)
        self.treecode = np.array([treecode_from_ijk(c, level, self.jmax, dim) for c in self.ijk],
                                 np.int64)
        x0dx = np.zeros((self.n, 6), np.float64)
        for k, c in enumerate(self.ijk):
            x0, dx = spacing_origin(c, level, self.domain, self.bs, dim)
            x0dx[k, :3] = x0
            x0dx[k, 3:] = dx
        self.x0dx = x0dx

    def block_at(self):
        """block_at[(k*nb + j)*nb + i] = hvy id (used by the oracle)."""
        return np.arange(self.n, dtype=np.int32)

    def hvy_of(self, i, j, k=0):
        nb = self.nb
        i, j, k = i % nb, j % nb, (k % nb if self.dim == 3 else 0)
        return (k * nb + j) * nb + i

    def same_level_patches(self, gs):
        """Same-level ghost patches of width gs on the periodic grid: the
        relations 1..8 (2D) / 1..26 (3D) of neighborhood.f90 with lvl_diff=0,
        boxes as ghosts_setup_patches builds them for g_minus = g_plus = gs
        (module_mpi.f90:318-396, calc_data_bounds.f90:46-189)."""
        g, bs, dim = self.g, self.bs, self.dim
        offs = [o for o in itertools.product((-1, 0, 1), repeat=dim) if any(o)]
        out = np.zeros(self.n * len(offs), dtype=PATCH_DTYPE)
        q = 0
        for hid, c in enumerate(self.ijk):
            for o in offs:
                o3 = list(o) + [0] * (3 - dim)
                nid = self.hvy_of(c[0] + o3[0], c[1] + o3[1], c[2] + o3[2])
                rlo, slo, n = [0, 0, 0], [0, 0, 0], [1, 1, 1]
                for d in range(dim):
                    if o3[d] == -1:
                        rlo[d], slo[d], n[d] = g - gs, g + bs[d] - gs, gs
                    elif o3[d] == 0:
                        rlo[d], slo[d], n[d] = g, g, bs[d]
                    else:
                        rlo[d], slo[d], n[d] = g + bs[d], g, gs
                out[q] = (hid, nid, rlo, slo, n, WH_PATCH_COPY, 1, 0)
                q += 1
        return out


def equidistant_treecodes(level: int, dim: int, jmax: int):
    """Treecodes and levels of createEquidistantGrid_tree at ``level``
    (createEquidistantGrid_tree.f90), x fastest."""
    n1 = 2 ** level
    tcs = []
    for k in range(n1 ** dim):
        ijk = [k % n1, (k // n1) % n1, (k // (n1 * n1)) if dim == 3 else 0]
        tcs.append(treecode_from_ijk(ijk, level, jmax, dim))
    return np.asarray(tcs, np.int64), np.full(len(tcs), level, np.int32)


def refined_treecodes(treecodes, levels, jmax_tc: int, dim: int, level_max: int):
    """The block list after refine_tree "everywhere" (refine_tree.f90 with
    respectJmaxJmin_tree): every block below level_max is replaced by its
    2^dim daughters, digit d at the daughter's level (refinementExecute.f90)."""
    tc_out, lv_out = [], []
    for tc, lv in zip(treecodes, levels):
        tc, lv = int(tc), int(lv)
        if lv >= level_max:
            tc_out.append(tc)
            lv_out.append(lv)
            continue
        for d in range(1 << dim):
            tc_out.append(tc | (d << ((jmax_tc - lv - 1) * dim)))
            lv_out.append(lv + 1)
    return np.asarray(tc_out, np.int64), np.asarray(lv_out, np.int32)
