"""CDF wavelets and adapt_tree over the C ABI (wh_wavelet_setup, wh_*_blocks,
wh_apply_boxes, wh_run_patches, wh_adapt_tree, wh_sync_tree).

Mirrors the reference's setup_wavelet (LIB/WAVELETS/module_wavelets.f90:1031-1430),
the per-block transforms of wavelet_decomposition_reconstruction.f90 and
adapt_tree (LIB/MESH/adapt_tree.f90:11-255).  There is no CPU fallback: every
call goes to the HIP library.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib as L

_TAPS = 33
_vp = ctypes.c_void_p
_BOUND = False

WH_BOX_ZERO_WC, WH_BOX_COPY_SC = 0, 1
WH_IND_EVERYWHERE, WH_IND_THRESHOLD, WH_IND_EXTERNAL = 0, 1, 2
WH_NORM_LINFTY, WH_NORM_L1, WH_NORM_L2, WH_NORM_H1 = 0, 1, 2, 3


class WaveletC(ctypes.Structure):
    _fields_ = [("X", ctypes.c_int), ("Y", ctypes.c_int), ("lifted", ctypes.c_int),
                ("specialised", ctypes.c_int),
                ("hd_lo", ctypes.c_int), ("hd_hi", ctypes.c_int), ("gd_lo", ctypes.c_int),
                ("gd_hi", ctypes.c_int), ("hr_lo", ctypes.c_int), ("hr_hi", ctypes.c_int),
                ("gr_lo", ctypes.c_int), ("gr_hi", ctypes.c_int),
                ("HD", ctypes.c_double * _TAPS), ("GD", ctypes.c_double * _TAPS),
                ("HR", ctypes.c_double * _TAPS), ("GR", ctypes.c_double * _TAPS),
                ("g", ctypes.c_int), ("g_rhs", ctypes.c_int), ("order", ctypes.c_int),
                ("Nscl", ctypes.c_int), ("Nscr", ctypes.c_int), ("Nwcl", ctypes.c_int),
                ("Nwcr", ctypes.c_int), ("Nreconl", ctypes.c_int), ("Nreconr", ctypes.c_int),
                ("Ndep2l", ctypes.c_int), ("Ndep2r", ctypes.c_int)]

    def filt(self, name):
        lo, hi = getattr(self, name.lower() + "_lo"), getattr(self, name.lower() + "_hi")
        return lo, np.array([getattr(self, name)[i] for i in range(hi - lo + 1)])


class BoxC(ctypes.Structure):
    _fields_ = [("hvy", ctypes.c_int32), ("lo", ctypes.c_int32 * 3), ("hi", ctypes.c_int32 * 3),
                ("op", ctypes.c_int32)]


class MaskParamsC(ctypes.Structure):
    _fields_ = [("geometry", ctypes.c_int), ("x_cntr", ctypes.c_double * 3), ("R_cyl", ctypes.c_double),
                ("smoothing_width", ctypes.c_double), ("smoothing_safety", ctypes.c_double),
                ("use_sponge", ctypes.c_int), ("sponge_type", ctypes.c_int),
                ("L_sponge", ctypes.c_double), ("p_sponge", ctypes.c_double),
                ("domain", ctypes.c_double * 3)]


class AdaptParamsC(ctypes.Structure):
    _fields_ = [("jmin", ctypes.c_int), ("jmax", ctypes.c_int), ("indicator", ctypes.c_int),
                ("use_ce", ctypes.c_int), ("force_maxlevel_dealiasing", ctypes.c_int),
                ("eps", ctypes.c_double), ("norm", ctypes.c_double * 8),
                ("groups", ctypes.c_int32 * 8), ("eps_normalized", ctypes.c_int),
                ("threshold_mask", ctypes.c_int), ("security_zone", ctypes.c_int),
                ("domain", ctypes.c_double * 3), ("mask", MaskParamsC), ("decompose_branch", ctypes.c_int),
                ("eps_norm", ctypes.c_int)]


def mask_params(geometry=0, x_cntr=(0.0, 0.0, 0.0), R_cyl=0.5, smoothing_width=0.0, smoothing_safety=0.0,
                use_sponge=False, sponge_type=0, L_sponge=0.0, p_sponge=20.0,
                domain=(1.0, 1.0, 1.0)) -> MaskParamsC:
    """The [VPM]/[Sponge] settings create_mask_2D_ACM reads (module_ACM.f90:300-481):
    geometry 1 = cylinder, sponge_type 0 = rect, 1 = p-norm."""
    m = MaskParamsC()
    m.geometry = geometry
    for i in range(3):
        m.x_cntr[i] = x_cntr[i] if i < len(x_cntr) else 0.0
        m.domain[i] = domain[i] if i < len(domain) else 1.0
    m.R_cyl, m.smoothing_width, m.smoothing_safety = R_cyl, smoothing_width, smoothing_safety
    m.use_sponge, m.sponge_type = int(use_sponge), sponge_type
    m.L_sponge, m.p_sponge = L_sponge, p_sponge
    return m


def create_mask(pool, mp: MaskParamsC, ids=None, x0dx=None):
    """The penalisation mask of the listed blocks (all active blocks when ids
    is None) into WH_MASK."""
    _bind()
    if ids is None:
        L.check(L.load().wh_create_mask(pool.h, ctypes.byref(mp), None, -1, None))
        return
    a = _ids(ids)
    x = np.ascontiguousarray(x0dx, np.float64)
    L.check(L.load().wh_create_mask(pool.h, ctypes.byref(mp), a.ctypes.data, a.size, x.ctypes.data))


BOX_DTYPE = np.dtype([("hvy", np.int32), ("lo", np.int32, 3), ("hi", np.int32, 3), ("op", np.int32)])


def _bind():
    global _BOUND
    if _BOUND:
        return
    lib = L.load()
    ip = ctypes.c_int
    sig = {
        "wh_wavelet_setup": [ctypes.c_char_p, ip, ctypes.POINTER(WaveletC)],
        "wh_pool_set_wavelet": [_vp, ctypes.POINTER(WaveletC)],
        "wh_fwt_blocks": [_vp, ip, ip, _vp, ip],
        "wh_iwt_blocks": [_vp, ip, ip, _vp, ip],
        "wh_filter_blocks": [_vp, ip, ip, _vp, ip],
        "wh_copy_blocks": [_vp, ip, ip, _vp, ip],
        "wh_detail_blocks": [_vp, ip, _vp, ip, _vp, _vp],
        "wh_detail_blocks_norm": [_vp, ip, _vp, _vp, ip, _vp, ip, _vp],
        "wh_component_norm": [_vp, ip, _vp, _vp, ip, _vp, ip, _vp, _vp],
        "wh_apply_boxes": [_vp, ip, ip, _vp, ip],
        "wh_run_patches": [_vp, ip, ip, ip, _vp, ip],
        "wh_adapt_tree": [_vp, _vp, ctypes.POINTER(AdaptParamsC)],
        "wh_adapt_tree_flags": [_vp, _vp, ctypes.POINTER(AdaptParamsC), _vp, ip],
        "wh_fwt_threshold": [_vp, ip, ip, _vp, ip, _vp, _vp, _vp, _vp],
        "wh_fwt_threshold_noalloc": [_vp, ip, ip, _vp, ip, _vp, _vp, _vp, _vp],
        "wh_sync_tree": [_vp, _vp, ip],
        "wh_sync_trees": [_vp, _vp, ip, ip],
        "wh_adapt_trees": [_vp, _vp, ip, _vp, _vp],
        "wh_refine_trees": [_vp, _vp, ip, _vp, ip, ip, _vp],
        "wh_create_masks": [_vp, _vp, ip, _vp, _vp],
        "wh_set_tables_from_forests": [_vp, _vp, ip, ip, ip, _vp, _vp],
        "wh_create_mask": [_vp, ctypes.POINTER(MaskParamsC), _vp, ip, _vp],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _BOUND = True


def setup(name: str, fd_max_size: int = 0) -> WaveletC:
    """setup_wavelet for "CDFXY"."""
    _bind()
    w = WaveletC()
    L.check(L.load().wh_wavelet_setup(name.encode(), fd_max_size, ctypes.byref(w)))
    return w


def set_wavelet(pool, w: WaveletC):
    _bind()
    L.check(L.load().wh_pool_set_wavelet(pool.h, ctypes.byref(w)))


def _ids(ids):
    return np.ascontiguousarray(ids, np.int32)


def _blocks_op(fn, pool, src, dst, ids):
    _bind()
    a = _ids(ids)
    L.check(getattr(L.load(), fn)(pool.h, src, dst, a.ctypes.data if a.size else None, a.size))


def fwt_blocks(pool, src, dst, ids):
    _blocks_op("wh_fwt_blocks", pool, src, dst, ids)


def iwt_blocks(pool, src, dst, ids):
    _blocks_op("wh_iwt_blocks", pool, src, dst, ids)


def filter_blocks(pool, src, dst, ids):
    _blocks_op("wh_filter_blocks", pool, src, dst, ids)


def copy_blocks(pool, src, dst, ids):
    _blocks_op("wh_copy_blocks", pool, src, dst, ids)


def detail_blocks(pool, array, ids, groups, neq):
    _bind()
    a = _ids(ids)
    g = np.ascontiguousarray(groups, np.int32)
    out = np.zeros((a.size, neq))
    L.check(L.load().wh_detail_blocks(pool.h, array, a.ctypes.data if a.size else None, a.size,
                                      g.ctypes.data, out.ctypes.data))
    return out


def detail_blocks_norm(pool, array, ids, levels, groups, neq, eps_norm):
    """threshold_block's detail with wavelet_renorm_block of eps_norm
    (WH_NORM_*) for blocks on the given levels."""
    _bind()
    a = _ids(ids)
    lv = np.ascontiguousarray(levels, np.int32)
    g = np.ascontiguousarray(groups, np.int32)
    assert lv.size == a.size
    out = np.zeros((a.size, neq))
    L.check(L.load().wh_detail_blocks_norm(pool.h, array, a.ctypes.data if a.size else None,
                                           lv.ctypes.data if a.size else None, a.size, g.ctypes.data, int(eps_norm),
                                           out.ctypes.data))
    return out


def component_norm(pool, array, ids, levels, groups, neq, eps_norm, domain):
    """componentWiseNorm_tree over the listed blocks (eps_norm WH_NORM_*)."""
    _bind()
    a = _ids(ids)
    lv = np.ascontiguousarray(levels, np.int32)
    g = np.ascontiguousarray(groups, np.int32)
    dom = np.ascontiguousarray(list(domain) + [1.0] * (3 - len(domain)), np.float64)
    assert lv.size == a.size
    out = np.zeros(neq)
    L.check(L.load().wh_component_norm(pool.h, array, a.ctypes.data if a.size else None,
                                       lv.ctypes.data if a.size else None, a.size, g.ctypes.data, int(eps_norm),
                                       dom.ctypes.data, out.ctypes.data))
    return out


def apply_boxes(pool, array, src_array, boxes):
    _bind()
    b = np.ascontiguousarray(boxes, BOX_DTYPE)
    L.check(L.load().wh_apply_boxes(pool.h, array, src_array, b.ctypes.data if b.size else None, b.size))


def run_patches(pool, array, patches, alt1=-1, alt2=-1):
    _bind()
    p = np.ascontiguousarray(patches, L.PATCH_DTYPE)
    L.check(L.load().wh_run_patches(pool.h, array, alt1, alt2, p.ctypes.data if p.size else None, p.size))


def adapt_params(jmin, jmax, indicator=WH_IND_THRESHOLD, use_ce=True, eps=1e-3, norm=None,
                 groups=None, force_maxlevel_dealiasing=False, neq=1, eps_normalized=False,
                 threshold_mask=False, security_zone=None, domain=(1.0, 1.0, 1.0),
                 mask=None, decompose_branch=0, eps_norm=0) -> AdaptParamsC:
    """adapt_tree settings; security_zone defaults to use_ce, as the
    reference's useSecurityZone defaults to the lifted-wavelet flag
    (ini_file_to_params.f90:543-546)."""
    p = AdaptParamsC()
    p.jmin, p.jmax, p.indicator = jmin, jmax, indicator
    p.decompose_branch = int(decompose_branch)
    p.eps_norm = int(eps_norm)
    p.use_ce = int(use_ce)
    p.eps_normalized = int(eps_normalized)
    p.threshold_mask = int(threshold_mask)
    p.security_zone = int(use_ce if security_zone is None else security_zone)
    for i in range(3):
        p.domain[i] = domain[i] if i < len(domain) else 1.0
    if mask is not None:
        p.mask = mask
    p.force_maxlevel_dealiasing = int(force_maxlevel_dealiasing)
    p.eps = eps
    for c in range(8):
        p.norm[c] = 1.0 if norm is None or c >= len(norm) else float(norm[c])
        p.groups[c] = 1 if groups is None or c >= len(groups) else int(groups[c])
    return p


def adapt_tree(pool, forest, params: AdaptParamsC):
    """adapt_tree on the forest's leaves in WH_BLOCK; the forest's block list
    becomes the new leaf list (re-attach the pool's tables afterwards)."""
    _bind()
    L.check(L.load().wh_adapt_tree(pool.h, forest.h, ctypes.byref(params)))


def adapt_tree_flags(pool, forest, params: AdaptParamsC, leaf_flags):
    """adapt_tree with the host's refinement flags (indicator "nothing
    (external)"): one flag per leaf in the forest's list order."""
    _bind()
    f = np.ascontiguousarray(leaf_flags, np.int32)
    L.check(L.load().wh_adapt_tree_flags(pool.h, forest.h, ctypes.byref(params), f.ctypes.data if f.size else None,
                                         f.size))


def fwt_threshold(pool, src, dst, ids, thresh, groups):
    """coarseningIndicator_block / threshold_block of the listed blocks:
    returns (flags (n,) in {-1, 0}, detail (n, neq))."""
    _bind()
    a = _ids(ids)
    th = np.ascontiguousarray(thresh, np.float64)
    g = np.ascontiguousarray(groups, np.int32)
    flags = np.zeros(a.size, np.int32)
    det = np.zeros((a.size, pool.neq))
    L.check(L.load().wh_fwt_threshold(pool.h, src, dst, a.ctypes.data if a.size else None, a.size, th.ctypes.data,
                                      g.ctypes.data, flags.ctypes.data, det.ctypes.data))
    return flags, det


def fwt_threshold_noalloc(pool, src, dst, ids, thresh, groups):
    """fwt_threshold through the entry point that allocates nothing (ids and
    details pass through the library's fixed device array; at most
    WH_NOALLOC_MAX_BLOCKS blocks per call)."""
    _bind()
    a = _ids(ids)
    th = np.ascontiguousarray(thresh, np.float64)
    g = np.ascontiguousarray(groups, np.int32)
    flags = np.zeros(a.size, np.int32)
    det = np.zeros((a.size, pool.neq))
    L.check(L.load().wh_fwt_threshold_noalloc(pool.h, src, dst, a.ctypes.data if a.size else None, a.size,
                                              th.ctypes.data, g.ctypes.data, flags.ctypes.data, det.ctypes.data))
    return flags, det


def sync_tree(pool, forest, array=L.WH_BLOCK):
    """sync_ghosts_tree with filtered restriction for lifted wavelets."""
    _bind()
    L.check(L.load().wh_sync_tree(pool.h, forest.h, array))


def sync_trees(pool, forests, array=L.WH_BLOCK):
    """sync_tree of several forests sharing the pool, one launch per
    operation for all of them."""
    _bind()
    if not forests:
        return
    arr = (ctypes.c_void_p * len(forests))(*[f.h.value for f in forests])
    L.check(L.load().wh_sync_trees(pool.h, ctypes.cast(arr, ctypes.c_void_p), len(forests), array))


def adapt_trees(pool, forests, params):
    """adapt_tree of several forests sharing the pool (params: one
    AdaptParamsC per forest), their device operations launched together.
    Returns one (code, message) per forest: code 0, or the error that forest
    alone hit (its block list is then unchanged).  An error of an operation
    shared by all forests raises."""
    _bind()
    n = len(forests)
    if n == 0:
        return []
    arr = (ctypes.c_void_p * n)(*[f.h.value for f in forests])
    prm = (AdaptParamsC * n)(*params)
    st = (ctypes.c_int * n)()
    L.check(L.load().wh_adapt_trees(pool.h, ctypes.cast(arr, ctypes.c_void_p), n, ctypes.cast(prm, ctypes.c_void_p),
                                    ctypes.cast(st, ctypes.c_void_p)))
    msg = L.load().wh_last_error().decode(errors="replace") if any(st) else ""
    return [(int(c), msg if c else "") for c in st]


def _status(st):
    msg = L.load().wh_last_error().decode(errors="replace") if any(st) else ""
    return [(int(c), msg if c else "") for c in st]


def refine_trees(pool, forests, jmax, order, array=L.WH_BLOCK):
    """refine_tree "everywhere" below jmax[i] for several forests sharing the
    pool, one refinement launch for all.  Returns one (code, message) per
    forest: code WH_E_NOMEM leaves a forest without the free slots for all its
    daughters at once unchanged."""
    _bind()
    n = len(forests)
    if n == 0:
        return []
    arr = (ctypes.c_void_p * n)(*[f.h.value for f in forests])
    jm = (ctypes.c_int32 * n)(*[int(j) for j in jmax])
    st = (ctypes.c_int * n)()
    L.check(L.load().wh_refine_trees(pool.h, ctypes.cast(arr, ctypes.c_void_p), n, ctypes.cast(jm, ctypes.c_void_p),
                                     order, array, ctypes.cast(st, ctypes.c_void_p)))
    return _status(st)


def create_masks(pool, forests, params):
    """createMask_tree on the leaves of several forests sharing the pool, each
    with its own MaskParamsC.  Returns one (code, message) per forest."""
    _bind()
    n = len(forests)
    if n == 0:
        return []
    arr = (ctypes.c_void_p * n)(*[f.h.value for f in forests])
    prm = (MaskParamsC * n)(*params)
    st = (ctypes.c_int * n)()
    L.check(L.load().wh_create_masks(pool.h, ctypes.cast(arr, ctypes.c_void_p), n, ctypes.cast(prm, ctypes.c_void_p),
                                     ctypes.cast(st, ctypes.c_void_p)))
    return _status(st)


def set_tables_from_forests(pool, forests, g_rhs, order):
    """The RK step's tables of the pool (active list, origins/spacings, one
    block group per forest, RHS ghost patches) from several forests' leaves.
    Returns ([(code, message)] per forest, [leaves in the tables] per forest);
    a forest with a non-zero code is left out."""
    _bind()
    n = len(forests)
    arr = (ctypes.c_void_p * max(n, 1))(*[f.h.value for f in forests])
    st = (ctypes.c_int * max(n, 1))()
    nb = (ctypes.c_int * max(n, 1))()
    L.check(L.load().wh_set_tables_from_forests(pool.h, ctypes.cast(arr, ctypes.c_void_p), n, g_rhs, order,
                                                ctypes.cast(st, ctypes.c_void_p), ctypes.cast(nb, ctypes.c_void_p)))
    pool.ngroups = sum(1 for c in st[:n] if c == 0)  # one block group per forest taken
    return _status(st[:n]), [int(x) for x in nb[:n]]

