"""ctypes bindings of the oracle (liboracle.so, built from oracle/*.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, never by wabbit_amd.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class OrAcmParams(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("discretization", ctypes.c_int),
        ("skew_symmetry", ctypes.c_int), ("penalization", ctypes.c_int),
        ("use_sponge", ctypes.c_int),
        ("c_0", ctypes.c_double), ("nu", ctypes.c_double), ("gamma_p", ctypes.c_double),
        ("C_eta", ctypes.c_double), ("C_sponge", ctypes.c_double),
        ("u_mean_set", ctypes.c_double * 3),
        ("CFL", ctypes.c_double), ("CFL_nu", ctypes.c_double), ("CFL_eta", ctypes.c_double),
        ("plant", ctypes.c_int),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise OSError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        L.or_rhs_acm_block.argtypes = [ctypes.POINTER(OrAcmParams), _ip, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.or_dt_block.argtypes = [ctypes.POINTER(OrAcmParams), _ip, ctypes.c_int, _dp, _dp]
        L.or_dt_block.restype = ctypes.c_double
        L.or_maxabs_block.argtypes = [_ip, ctypes.c_int, ctypes.c_int, _dp]
        L.or_maxabs_block.restype = ctypes.c_double
        L.or_sync_equidistant.argtypes = [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          _ip, ctypes.c_int, _dp, ctypes.c_int]
        L.or_rk_step_equidistant.argtypes = [ctypes.POINTER(OrAcmParams), _ip, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int, _dp, _dp, _dp,
                                             _dp, _dp, ctypes.c_int, ctypes.c_double]
        L.or_dt_min.argtypes = [ctypes.POINTER(OrAcmParams), _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                _dp, _dp]
        L.or_dt_min.restype = ctypes.c_double
        L.or_sync_leaf.argtypes = [ctypes.POINTER(OrForest), _dp, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int]
        L.or_refine_block.argtypes = [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, _dp,
                                      ctypes.POINTER(_dp), ctypes.c_int]
        L.or_prediction.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int]
        _LIB = L
    return _LIB


class OrForest(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("jmax", ctypes.c_int), ("g", ctypes.c_int),
                ("bs", ctypes.c_int * 3), ("n", ctypes.c_int),
                ("tc", ctypes.POINTER(ctypes.c_int64)), ("level", ctypes.POINTER(ctypes.c_int))]


class LeafForest:
    """A leaf forest for the oracle: block b <-> row b of the heavy array."""

    def __init__(self, dim, jmax, bs, g, treecodes, levels):
        self.dim, self.jmax, self.g = dim, jmax, g
        self.bs = list(bs) + [1] * (3 - len(bs))
        self.tc = np.ascontiguousarray(treecodes, np.int64)
        self.level = np.ascontiguousarray(levels, np.int32)
        f = OrForest()
        f.dim, f.jmax, f.g = dim, jmax, g
        for i in range(3):
            f.bs[i] = self.bs[i]
        f.n = self.tc.shape[0]
        f.tc = self.tc.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        f.level = self.level.ctypes.data_as(_ip)
        self.c = f

    def sync(self, hvy, gm, gp, order):
        """Ghost fill of depth gm/gp in place; hvy (n, neq, nz, ny, nx)."""
        assert hvy.flags.c_contiguous and hvy.dtype == np.float64
        rc = lib().or_sync_leaf(ctypes.byref(self.c), _d(hvy), hvy.shape[1], gm, gp, order)
        if rc:
            raise RuntimeError(f"or_sync_leaf failed ({rc})")

    def refine_all(self, hvy, order):
        """refineBlock of every block below jmax; returns (treecodes, levels, hvy)."""
        dim, g, bs = self.dim, self.g, self.bs
        out_tc, out_lev, out = [], [], []
        nd = 1 << dim
        bsa = np.asarray(bs, np.int32)
        for b in range(self.tc.shape[0]):
            lev = int(self.level[b])
            if lev >= self.jmax:
                out_tc.append(int(self.tc[b]))
                out_lev.append(lev)
                out.append(hvy[b].copy())
                continue
            daughters = [np.zeros_like(hvy[b]) for _ in range(nd)]
            arr = (_dp * nd)(*[_d(d) for d in daughters])
            m = np.ascontiguousarray(hvy[b])
            lib().or_refine_block(dim, _i(bsa), g, hvy.shape[1], _d(m), arr, order)
            for k in range(nd):
                out_tc.append(int(self.tc[b]) | (k << ((self.jmax - lev - 1) * dim)))
                out_lev.append(lev + 1)
                out.append(daughters[k])
        return np.array(out_tc, np.int64), np.array(out_lev, np.int32), np.ascontiguousarray(np.stack(out))


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_ip)


def make_params(acm) -> OrAcmParams:
    """From a wabbit_amd.params.ACMParams-like object (duck typed)."""
    p = OrAcmParams()
    p.dim = acm.dim
    p.discretization = acm.scheme
    p.skew_symmetry = int(acm.skew_symmetry)
    p.penalization = int(acm.penalization)
    p.use_sponge = int(acm.use_sponge)
    p.c_0, p.nu, p.gamma_p = acm.c_0, acm.nu, acm.gamma_p
    p.C_eta, p.C_sponge = acm.C_eta, acm.C_sponge
    for i in range(3):
        p.u_mean_set[i] = acm.u_mean_set[i]
    p.CFL, p.CFL_nu, p.CFL_eta = acm.CFL, acm.CFL_nu, acm.CFL_eta
    return p


def rhs_block(p: OrAcmParams, Bs, g, dx, u, mask=None):
    u = np.ascontiguousarray(u, dtype=np.float64)
    rhs = np.zeros_like(u)
    bs = np.asarray(list(Bs) + [1] * (3 - len(Bs)), dtype=np.int32)
    dxa = np.asarray(list(dx) + [0.0] * (3 - len(dx)), dtype=np.float64)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.float64)
    lib().or_rhs_acm_block(ctypes.byref(p), _i(bs), g, _d(dxa), _d(u), _d(rhs), _d(m))
    return rhs


def dt_block(p: OrAcmParams, Bs, g, dx, u):
    bs = np.asarray(list(Bs) + [1] * (3 - len(Bs)), dtype=np.int32)
    dxa = np.asarray(list(dx) + [0.0] * (3 - len(dx)), dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().or_dt_block(ctypes.byref(p), _i(bs), g, _d(dxa), _d(u))


class EquidistantRun:
    """Oracle state of an equidistant periodic forest (all blocks at level J).

    hvy: (Nb, neq, nz, ny, nx) float64 in the reference's block layout."""

    def __init__(self, p: OrAcmParams, Bs, g, g_rhs, neq, J, domain, hvy, block_at, mask=None,
                 butcher=None):
        self.p = p
        self.dim = p.dim
        self.Bs = np.asarray(list(Bs) + [1] * (3 - len(Bs)), dtype=np.int32)
        self.g, self.g_rhs, self.neq = g, g_rhs, neq
        self.nb = 2 ** J
        self.hvy = np.ascontiguousarray(hvy, dtype=np.float64)
        self.nblocks = self.hvy.shape[0]
        self.block_at = np.ascontiguousarray(block_at, dtype=np.int32)
        dx = [2.0 ** (-J) * domain[d] / self.Bs[d] for d in range(self.dim)] + [0.0] * (3 - self.dim)
        self.dxs = np.tile(np.asarray(dx, np.float64), (self.nblocks, 1))
        self.mask = None if mask is None else np.ascontiguousarray(mask, dtype=np.float64)
        if butcher is None:
            butcher = np.array([[0, 0, 0, 0, 0], [0.5, 0.5, 0, 0, 0], [0.5, 0, 0.5, 0, 0],
                                [1, 0, 0, 1, 0], [0, 1 / 6, 1 / 3, 1 / 3, 1 / 6]], np.float64)
        self.butcher = np.ascontiguousarray(butcher, np.float64)
        self.work = np.zeros((self.butcher.shape[0],) + self.hvy.shape, np.float64)

    def dt(self):
        return lib().or_dt_min(ctypes.byref(self.p), _i(self.Bs), self.g, self.neq, self.nblocks,
                               _d(self.dxs), _d(self.hvy))

    def sync(self, gs=None):
        lib().or_sync_equidistant(self.dim, _i(self.Bs), self.g, self.neq, self.nb, _i(self.block_at),
                                  self.nblocks, _d(self.hvy), self.g_rhs if gs is None else gs)

    def step(self, dt):
        lib().or_rk_step_equidistant(ctypes.byref(self.p), _i(self.Bs), self.g, self.g_rhs, self.neq,
                                     self.nb, _i(self.block_at), self.nblocks, _d(self.dxs),
                                     _d(self.hvy), _d(self.work), _d(self.mask), _d(self.butcher),
                                     self.butcher.shape[0], dt)


# ---- wavelets and adapt_tree (wavelet.c, tree.c) ----------------------------
_WMAX = 16


class OrWavelet(ctypes.Structure):
    _fields_ = [("X", ctypes.c_int), ("Y", ctypes.c_int), ("lifted", ctypes.c_int),
                ("hd_lo", ctypes.c_int), ("hd_hi", ctypes.c_int), ("gd_lo", ctypes.c_int),
                ("gd_hi", ctypes.c_int), ("hr_lo", ctypes.c_int), ("hr_hi", ctypes.c_int),
                ("gr_lo", ctypes.c_int), ("gr_hi", ctypes.c_int),
                ("HD", ctypes.c_double * (2 * _WMAX + 1)), ("GD", ctypes.c_double * (2 * _WMAX + 1)),
                ("HR", ctypes.c_double * (2 * _WMAX + 1)), ("GR", ctypes.c_double * (2 * _WMAX + 1)),
                ("g", ctypes.c_int), ("g_rhs", ctypes.c_int), ("order", ctypes.c_int),
                ("Nscl", ctypes.c_int), ("Nscr", ctypes.c_int), ("Nwcl", ctypes.c_int),
                ("Nwcr", ctypes.c_int), ("Nreconl", ctypes.c_int), ("Nreconr", ctypes.c_int),
                ("Ndep2l", ctypes.c_int), ("Ndep2r", ctypes.c_int), ("specialised", ctypes.c_int)]

    def filt(self, name):
        lo, hi = getattr(self, name.lower() + "_lo"), getattr(self, name.lower() + "_hi")
        return lo, np.array([getattr(self, name)[i] for i in range(hi - lo + 1)])


class OrMaskParams(ctypes.Structure):
    _fields_ = [("geometry", ctypes.c_int), ("x_cntr", ctypes.c_double * 3), ("R_cyl", ctypes.c_double),
                ("smoothing_width", ctypes.c_double), ("smoothing_safety", ctypes.c_double),
                ("use_sponge", ctypes.c_int), ("sponge_type", ctypes.c_int),
                ("L_sponge", ctypes.c_double), ("p_sponge", ctypes.c_double),
                ("domain", ctypes.c_double * 3)]


class OrAdaptParams(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("jmin", ctypes.c_int), ("jmax", ctypes.c_int),
                ("g", ctypes.c_int), ("neq", ctypes.c_int), ("bs", ctypes.c_int * 3),
                ("indicator", ctypes.c_int), ("use_ce", ctypes.c_int),
                ("force_maxlevel_dealiasing", ctypes.c_int), ("eps", ctypes.c_double),
                ("norm", ctypes.c_double * 8), ("groups", ctypes.c_int * 8),
                ("eps_normalized", ctypes.c_int), ("threshold_mask", ctypes.c_int),
                ("security_zone", ctypes.c_int), ("domain", ctypes.c_double * 3),
                ("mask", OrMaskParams), ("decompose_branch", ctypes.c_int), ("plant", ctypes.c_int),
                ("eps_norm", ctypes.c_int)]

# params%eps_norm
NORM_LINFTY, NORM_L1, NORM_L2, NORM_H1 = 0, 1, 2, 3


def mask_params(geometry=0, x_cntr=(0, 0, 0), R_cyl=0.5, smoothing_width=0.0, smoothing_safety=0.0,
                use_sponge=False, sponge_type=0, L_sponge=0.0, p_sponge=20.0, domain=(1, 1, 1)):
    m = OrMaskParams()
    m.geometry = geometry
    for i in range(3):
        m.x_cntr[i] = x_cntr[i]
        m.domain[i] = domain[i]
    m.R_cyl, m.smoothing_width, m.smoothing_safety = R_cyl, smoothing_width, smoothing_safety
    m.use_sponge, m.sponge_type = int(use_sponge), sponge_type
    m.L_sponge, m.p_sponge = L_sponge, p_sponge
    return m


def mask_block(mp, dim, bs, g, x0, dx):
    _wlib()
    nx = bs[0] + 2 * g
    nz = bs[2] + 2 * g if dim == 3 else 1
    out = np.zeros((6, nz, bs[1] + 2 * g, nx))
    x0a = np.asarray(list(x0) + [0.0] * (3 - len(x0)), np.float64)
    dxa = np.asarray(list(dx) + [0.0] * (3 - len(dx)), np.float64)
    lib().or_mask_block(ctypes.byref(mp), dim, _i(_blk(dim, bs)), g, _d(x0a), _d(dxa), _d(out))
    return out


def _wlib():
    L = lib()
    if not getattr(L, "_wavelet_bound", False):
        vp = ctypes.c_void_p
        L.or_wavelet_setup.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(OrWavelet)]
        L.or_wavelet_setup.restype = ctypes.c_int
        for fn in ("or_fwt_block", "or_iwt_block", "or_filter_block"):
            getattr(L, fn).argtypes = [ctypes.POINTER(OrWavelet), ctypes.c_int, _ip, ctypes.c_int,
                                       ctypes.c_int, _dp, _dp]
        L.or_detail_block.argtypes = [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, _ip, _dp, _dp]
        L.or_detail_block_norm.argtypes = [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, _ip, _dp, ctypes.c_int,
                                           ctypes.c_int, _dp]
        L.or_set_norm_plant.argtypes = [ctypes.c_int]
        L.or_renorm_value.argtypes = [ctypes.c_int] * 7 + [ctypes.c_double]
        L.or_renorm_value.restype = ctypes.c_double
        L.or_component_norm.argtypes = [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int,
                                        ctypes.c_int, ctypes.POINTER(_dp), _dp, _dp]
        L.or_tree_create.argtypes = [ctypes.POINTER(OrWavelet), ctypes.POINTER(OrAdaptParams), ctypes.c_int,
                                     vp, _ip, _dp]
        L.or_tree_create.restype = vp
        L.or_tree_destroy.argtypes = [vp]
        L.or_tree_size.argtypes = [vp]
        L.or_tree_get.argtypes = [vp, vp, _ip, _dp]
        for fn in ("or_tree_sync_full", "or_tree_refine_everywhere", "or_tree_adapt"):
            getattr(L, fn).argtypes = [vp]
            getattr(L, fn).restype = ctypes.c_int
        L.or_tree_adapt_flags.argtypes = [vp, _ip]
        L.or_tree_adapt_flags.restype = ctypes.c_int
        L.or_mask_block.argtypes = [ctypes.POINTER(OrMaskParams), ctypes.c_int, _ip, ctypes.c_int, _dp, _dp, _dp]
        L.or_tree_dt.argtypes = [vp, ctypes.POINTER(OrAcmParams)]
        L.or_tree_dt.restype = ctypes.c_double
        L.or_tree_rk_step.argtypes = [vp, ctypes.POINTER(OrAcmParams), ctypes.POINTER(OrMaskParams), _dp,
                                      ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.or_tree_rk_step.restype = ctypes.c_int
        L.or_tree_get_mask.argtypes = [vp, ctypes.POINTER(OrMaskParams), _dp]
        L._wavelet_bound = True
    return L


def wavelet(name, fd_max_size=0) -> OrWavelet:
    w = OrWavelet()
    if _wlib().or_wavelet_setup(name.encode(), fd_max_size, ctypes.byref(w)):
        raise ValueError(f"unsupported wavelet {name}")
    return w


def _blk(dim, bs):
    return np.asarray(list(bs) + [1] * (3 - len(bs)), np.int32)


def fwt_block(w, dim, bs, g, src, dst=None):
    src = np.ascontiguousarray(src, np.float64)
    dst = src.copy() if dst is None else np.ascontiguousarray(dst, np.float64).copy()
    _wlib().or_fwt_block(ctypes.byref(w), dim, _i(_blk(dim, bs)), g, src.shape[0], _d(src), _d(dst))
    return dst


def iwt_block(w, dim, bs, g, src, dst=None):
    src = np.ascontiguousarray(src, np.float64)
    dst = src.copy() if dst is None else np.ascontiguousarray(dst, np.float64).copy()
    _wlib().or_iwt_block(ctypes.byref(w), dim, _i(_blk(dim, bs)), g, src.shape[0], _d(src), _d(dst))
    return dst


def filter_block(w, dim, bs, g, src):
    src = np.ascontiguousarray(src, np.float64)
    dst = np.empty_like(src)
    _wlib().or_filter_block(ctypes.byref(w), dim, _i(_blk(dim, bs)), g, src.shape[0], _d(src), _d(dst))
    return dst


def detail_block(dim, bs, g, u, groups):
    u = np.ascontiguousarray(u, np.float64)
    gr = np.ascontiguousarray(groups, np.int32)
    out = np.zeros(u.shape[0])
    _wlib().or_detail_block(dim, _i(_blk(dim, bs)), g, u.shape[0], _i(gr), _d(u), _d(out))
    return out


def detail_block_norm(dim, bs, g, u, groups, level, eps_norm):
    """threshold_block's detail with wavelet_renorm_block of eps_norm for a
    block on `level` (threshold_block.f90:69-124, module_wavelets.f90:1848-1960)."""
    u = np.ascontiguousarray(u, np.float64)
    gr = np.ascontiguousarray(groups, np.int32)
    out = np.zeros(u.shape[0])
    _wlib().or_detail_block_norm(dim, _i(_blk(dim, bs)), g, u.shape[0], _i(gr), _d(u), int(level), int(eps_norm),
                                 _d(out))
    return out


def renorm_value(eps_norm, dim, level, ix, iy, iz, g, v):
    """wavelet_renorm_block for one value at 0-based point (ix, iy, iz)."""
    return _wlib().or_renorm_value(int(eps_norm), dim, int(level), ix, iy, iz, g, float(v))


def set_norm_plant(bits):
    """Planted mistakes of the eps_norm functions (tests only; 0 restores)."""
    _wlib().or_set_norm_plant(int(bits))


def component_norm(dim, bs, g, blocks, dx, groups, eps_norm):
    """componentWiseNorm_tree over the given blocks (each (neq, nz, ny, nx)),
    dx[k] the spacing of block k (componentWiseNorm_tree.f90:1-200)."""
    blocks = [np.ascontiguousarray(b, np.float64) for b in blocks]
    neq = blocks[0].shape[0]
    ptrs = (_dp * len(blocks))(*[_d(b) for b in blocks])
    dxa = np.zeros((len(blocks), 3))
    dxa[:, :dim] = np.asarray(dx, np.float64).reshape(len(blocks), -1)[:, :dim]
    gr = np.ascontiguousarray(groups, np.int32)
    out = np.zeros(neq)
    _wlib().or_component_norm(dim, _i(_blk(dim, bs)), g, neq, _i(gr), int(eps_norm), len(blocks), ptrs,
                              _d(dxa), _d(out))
    return out


class Tree:
    """A periodic forest owned by the oracle (adapt_tree, syncs, refinement)."""

    def __init__(self, w, dim, jmin, jmax, bs, g, treecodes, levels, hvy, indicator=0, use_ce=None,
                 eps=1e-3, norm=None, groups=None, force_maxlevel_dealiasing=False, eps_normalized=False,
                 threshold_mask=False, security_zone=None, domain=(1.0, 1.0, 1.0), mask=None,
                 decompose_branch=0, plant=0, eps_norm=NORM_LINFTY):
        p = OrAdaptParams()
        p.decompose_branch = int(decompose_branch)
        p.eps_norm = int(eps_norm)
        p.plant = int(plant)
        p.dim, p.jmin, p.jmax, p.g = dim, jmin, jmax, g
        p.eps_normalized = int(eps_normalized)
        p.threshold_mask = int(threshold_mask)
        p.security_zone = int(w.lifted if security_zone is None else security_zone)
        for i in range(3):
            p.domain[i] = domain[i] if i < len(domain) else 1.0
        if mask is not None:
            p.mask = mask
        hvy = np.ascontiguousarray(hvy, np.float64)
        p.neq = hvy.shape[1]
        for i, b in enumerate(_blk(dim, bs)):
            p.bs[i] = int(b)
        p.indicator = indicator
        p.use_ce = int(w.lifted if use_ce is None else use_ce)
        p.force_maxlevel_dealiasing = int(force_maxlevel_dealiasing)
        p.eps = eps
        for c in range(p.neq):
            p.norm[c] = 1.0 if norm is None else norm[c]
            p.groups[c] = 1 if groups is None else groups[c]
        self.w, self.p = w, p
        self.dim, self.g, self.bs = dim, g, list(_blk(dim, bs))
        self.shape = hvy.shape[1:]
        tc = np.ascontiguousarray(treecodes, np.int64)
        lv = np.ascontiguousarray(levels, np.int32)
        self.h = _wlib().or_tree_create(ctypes.byref(w), ctypes.byref(p), tc.shape[0], tc.ctypes.data,
                                        _i(lv), _d(hvy))

    def close(self):
        if self.h:
            _wlib().or_tree_destroy(self.h)
            self.h = None

    __del__ = close

    def _ck(self, rc, what):
        if rc:
            raise RuntimeError(f"{what} failed ({rc})")

    def sync(self):
        self._ck(_wlib().or_tree_sync_full(self.h), "or_tree_sync_full")

    def refine_everywhere(self):
        self._ck(_wlib().or_tree_refine_everywhere(self.h), "or_tree_refine_everywhere")

    def adapt(self):
        self._ck(_wlib().or_tree_adapt(self.h), "or_tree_adapt")

    def adapt_flags(self, leaf_flags):
        """adapt_tree with external flags, one per leaf in get()'s order."""
        f = np.ascontiguousarray(leaf_flags, np.int32)
        self._ck(_wlib().or_tree_adapt_flags(self.h, _i(f)), "or_tree_adapt_flags")

    def size(self):
        return _wlib().or_tree_size(self.h)

    def get(self):
        """(treecodes, levels, data) sorted by (level, treecode)."""
        n = _wlib().or_tree_size(self.h)
        tc = np.zeros(n, np.int64)
        lv = np.zeros(n, np.int32)
        hv = np.zeros((n,) + tuple(self.shape))
        _wlib().or_tree_get(self.h, tc.ctypes.data, _i(lv), _d(hv))
        return tc, lv, hv

    def get_mask(self, mp):
        """Mask of the blocks, in get()'s order, shape (n, 6, nz, ny, nx)."""
        n = _wlib().or_tree_size(self.h)
        out = np.zeros((n, 6) + tuple(self.shape[1:]))
        _wlib().or_tree_get_mask(self.h, ctypes.byref(mp), _d(out))
        return out

    def dt(self, acm_p):
        return _wlib().or_tree_dt(self.h, ctypes.byref(acm_p))

    def rk_step(self, acm_p, mp, butcher, dt, g_rhs):
        bt = np.ascontiguousarray(butcher, np.float64)
        self._ck(_wlib().or_tree_rk_step(self.h, ctypes.byref(acm_p), ctypes.byref(mp) if mp is not None else None,
                                         _d(bt), bt.shape[0], dt, g_rhs), "or_tree_rk_step")
