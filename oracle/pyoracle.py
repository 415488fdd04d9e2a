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
