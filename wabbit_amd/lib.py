"""ctypes binding of libwabbit_hip.so (include/wabbit_hip.h).

The product path: every call below runs HIP kernels on the GPU.  There is no
CPU fallback -- if the library is not built, or no GPU is visible, the calls
raise.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libwabbit_hip.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "wabbit_hip.h")

WH_BLOCK, WH_TMP, WH_MASK, WH_WORK = 0, 1, 2, 16
WH_PATCH_COPY, WH_PATCH_RESTRICT, WH_PATCH_PREDICT = 0, 1, 2
WH_TUNE_K3D_YSPLIT, WH_TUNE_K2D_STAGE, WH_TUNE_K3D_ZSPLIT, WH_TUNE_K3D_WAVES, WH_TUNE_K3D_XCD = 1, 2, 3, 4, 5
WH_TUNE_GATHER = 6
ERRORS = {-1: "WH_E_ARG", -2: "WH_E_HIP", -3: "WH_E_STATE", -4: "WH_E_DIVERGED", -5: "WH_E_NOMEM"}
WH_OK, WH_E_ARG, WH_E_HIP, WH_E_STATE, WH_E_DIVERGED, WH_E_NOMEM = 0, -1, -2, -3, -4, -5


class WabbitHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class PoolDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("bs", ctypes.c_int * 3), ("g", ctypes.c_int),
                ("neq", ctypes.c_int), ("nb_max", ctypes.c_int), ("n_work_slots", ctypes.c_int),
                ("n_tmp_eqn", ctypes.c_int), ("n_mask", ctypes.c_int),
                ("mask_per_block", ctypes.c_int)]


class AcmParamsC(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("scheme", ctypes.c_int), ("skew_symmetry", ctypes.c_int),
                ("penalization", ctypes.c_int), ("use_sponge", ctypes.c_int),
                ("c_0", ctypes.c_double), ("nu", ctypes.c_double), ("gamma_p", ctypes.c_double),
                ("C_eta", ctypes.c_double), ("C_sponge", ctypes.c_double),
                ("u_mean_set", ctypes.c_double * 3), ("CFL", ctypes.c_double),
                ("CFL_nu", ctypes.c_double), ("CFL_eta", ctypes.c_double)]


class TimeParamsC(ctypes.Structure):
    _fields_ = [("time_max", ctypes.c_double), ("dt_fixed", ctypes.c_double),
                ("dt_max", ctypes.c_double), ("write_time", ctypes.c_double),
                ("write_time_first", ctypes.c_double), ("tsave_stats", ctypes.c_double),
                ("write_backup_time", ctypes.c_double), ("write_fixed_time", ctypes.c_int),
                ("write_backup_n", ctypes.c_int)]


# wh_patch as a numpy structured dtype (16 int32)
PATCH_DTYPE = np.dtype([("recv", "<i4"), ("send", "<i4"), ("recv_lo", "<i4", 3),
                        ("send_lo", "<i4", 3), ("n", "<i4", 3), ("op", "<i4"),
                        ("stage", "<i4"), ("aux", "<i4")])

_vp = ctypes.c_void_p
_LIB = None


def declared_symbols():
    """Function names declared in include/wabbit_hip.h."""
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s+\*?(wh_\w+)\s*\(", text, re.M)))


def load(path: str = LIB_PATH):
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise OSError(f"{path} is not built: run `python -m wabbit_amd.build` "
                      "(there is no CPU fallback)")
    L = ctypes.CDLL(path)
    L.wh_last_error.restype = ctypes.c_char_p
    pp = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "wh_pool_create": [pp, ctypes.POINTER(PoolDesc)],
        "wh_pool_destroy": [_vp],
        "wh_pool_block_bytes": [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)],
        "wh_pool_upload": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp],
        "wh_pool_download": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp],
        "wh_pool_device_ptr": [_vp, ctypes.c_int, pp],
        "wh_pool_fill": [_vp, ctypes.c_int, ctypes.c_double],
        "wh_synchronize": [_vp],
        "wh_set_active": [_vp, _vp, ctypes.c_int, _vp],
        "wh_set_patches": [_vp, ctypes.c_int, _vp, ctypes.c_int],
        "wh_sync_ghosts": [_vp, ctypes.c_int, ctypes.c_int],
        "wh_rhs_acm": [_vp, ctypes.POINTER(AcmParamsC), ctypes.c_double, ctypes.c_int, ctypes.c_int],
        "wh_divergence_guard": [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
        "wh_dt_reduce": [_vp, ctypes.POINTER(AcmParamsC), ctypes.POINTER(ctypes.c_double)],
        "wh_rk_combine": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp],
        "wh_rk_step": [_vp, ctypes.POINTER(AcmParamsC), _vp, ctypes.c_int, ctypes.c_double,
                       ctypes.c_double, ctypes.c_int],
        "wh_time_set": [_vp, ctypes.c_double, ctypes.c_int],
        "wh_time_get": [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                        ctypes.POINTER(ctypes.c_double)],
        "wh_time_fetch": [_vp],
        "wh_time_stage_kernel": [_vp, ctypes.POINTER(AcmParamsC), ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
        "wh_time_fetched": [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                            ctypes.POINTER(ctypes.c_double)],
        "wh_run_steps": [_vp, ctypes.POINTER(AcmParamsC), _vp, ctypes.c_int,
                         ctypes.POINTER(TimeParamsC), ctypes.c_int, ctypes.c_int],
        "wh_set_tuning": [_vp, ctypes.c_int, ctypes.c_int],
        "wh_get_tuning": [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
        "wh_autotune": [_vp, ctypes.POINTER(AcmParamsC), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                        ctypes.POINTER(ctypes.c_double)],
        "wh_kernel_timing_enable": [_vp, ctypes.c_int],
        "wh_kernel_timing": [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(ctypes.c_int)],
        "wh_set_block_groups": [_vp, _vp, ctypes.c_int, ctypes.c_int],
        "wh_dt_reduce_groups": [_vp, ctypes.POINTER(AcmParamsC), _vp],
        "wh_rk_step_groups": [_vp, ctypes.POINTER(AcmParamsC), _vp, ctypes.c_int, _vp, ctypes.c_int],
        "wh_set_group_params": [_vp, _vp, ctypes.c_int],
        "wh_batch_create": [pp, _vp, ctypes.c_int],
        "wh_batch_destroy": [_vp],
        "wh_batch_rk_step": [_vp, ctypes.POINTER(AcmParamsC), _vp, ctypes.c_int, _vp, ctypes.c_int],
        "wh_batch_time_stage_kernel": [_vp, ctypes.POINTER(AcmParamsC), _vp, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)],
        "wh_batch_synchronize": [_vp],
        "wh_batch_set_tuning": [_vp, ctypes.c_int, ctypes.c_int],
        "wh_batch_get_tuning": [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
        "wh_device_memory": [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _LIB = L
    return L


def device_memory():
    """(free, total) bytes of the current device, from the library's own HIP
    runtime (torch bundles another one; initialising that one first leaves
    this one without a device)."""
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    check(load().wh_device_memory(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


def check(rc):
    if rc != 0:
        raise WabbitHipError(rc, load().wh_last_error().decode(errors="replace"))
    return rc


def acm_params_c(acm) -> AcmParamsC:
    p = AcmParamsC()
    p.dim = acm.dim
    p.scheme = acm.scheme
    p.skew_symmetry = int(acm.skew_symmetry)
    p.penalization = int(acm.penalization)
    p.use_sponge = int(acm.use_sponge)
    p.c_0, p.nu, p.gamma_p = acm.c_0, acm.nu, acm.gamma_p
    p.C_eta, p.C_sponge = acm.C_eta, acm.C_sponge
    for i in range(3):
        p.u_mean_set[i] = acm.u_mean_set[i]
    p.CFL, p.CFL_nu, p.CFL_eta = acm.CFL, acm.CFL_nu, acm.CFL_eta
    return p


def time_params_c(tp) -> TimeParamsC:
    t = TimeParamsC()
    t.time_max, t.dt_fixed, t.dt_max = tp.time_max, tp.dt_fixed, tp.dt_max
    t.write_time, t.write_time_first = tp.write_time, tp.write_time_first
    t.tsave_stats, t.write_backup_time = tp.tsave_stats, tp.write_backup_time
    t.write_fixed_time = int(tp.write_method == "fixed_time")
    t.write_backup_n = tp.write_backup_n
    return t


class Pool:
    """A device block pool: the heavy-data arrays of one forest
    (allocate_forest.f90) resident in HBM."""

    def __init__(self, dim, bs, g, neq, nb_max, n_work_slots=5, n_tmp_eqn=0, n_mask=0,
                 mask_per_block=True):
        L = load()
        d = PoolDesc()
        d.dim = dim
        bs = list(bs) + [1] * (3 - len(bs))
        for i in range(3):
            d.bs[i] = int(bs[i])
        d.g, d.neq, d.nb_max = g, neq, nb_max
        d.n_work_slots, d.n_tmp_eqn, d.n_mask = n_work_slots, n_tmp_eqn, n_mask
        d.mask_per_block = int(mask_per_block)
        h = ctypes.c_void_p()
        check(L.wh_pool_create(ctypes.byref(h), ctypes.byref(d)))
        self.h = h
        self.dim, self.bs, self.g, self.neq, self.nb_max = dim, bs, g, neq, nb_max
        self.nx = bs[0] + 2 * g
        self.ny = bs[1] + 2 * g
        self.nz = bs[2] + 2 * g if dim == 3 else 1
        self.n_work_slots = n_work_slots
        self.n_tmp_eqn = n_tmp_eqn
        self.n_mask = n_mask
        self.mask_per_block = mask_per_block

    def close(self):
        if getattr(self, "h", None):
            load().wh_pool_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ncomp(self, array):
        """Components per block of a pool array."""
        if array == WH_TMP:
            return self.n_tmp_eqn
        if array == WH_MASK:
            return self.n_mask
        return self.neq

    def block_shape(self, ncomp=None):
        return (self.neq if ncomp is None else ncomp, self.nz, self.ny, self.nx)

    def upload(self, array, first, data):
        data = np.ascontiguousarray(data, dtype=np.float64)
        n = data.shape[0]
        if data.shape[1:] != self.block_shape(self.ncomp(array)):
            raise ValueError(f"upload: blocks of shape {data.shape[1:]}, the array holds "
                             f"{self.block_shape(self.ncomp(array))}")
        check(load().wh_pool_upload(self.h, array, first, n, data.ctypes.data))

    def download(self, array, first, count):
        out = np.empty((count,) + self.block_shape(self.ncomp(array)), np.float64)
        check(load().wh_pool_download(self.h, array, first, count, out.ctypes.data))
        return out

    def fill(self, array, value):
        check(load().wh_pool_fill(self.h, array, float(value)))

    def set_active(self, active, x0dx):
        a = np.ascontiguousarray(active, dtype=np.int32)
        x = np.ascontiguousarray(x0dx, dtype=np.float64).reshape(-1, 6)
        assert x.shape[0] == a.shape[0]
        check(load().wh_set_active(self.h, a.ctypes.data, a.shape[0], x.ctypes.data))

    def set_patches(self, sync_case, patches):
        p = np.ascontiguousarray(patches, dtype=PATCH_DTYPE)
        check(load().wh_set_patches(self.h, sync_case, p.ctypes.data if p.size else None, p.shape[0]))

    def sync(self, array=WH_BLOCK, sync_case=0):
        check(load().wh_sync_ghosts(self.h, array, sync_case))

    def rhs(self, acm_c, time, in_array, out_slot):
        check(load().wh_rhs_acm(self.h, ctypes.byref(acm_c), time, in_array, out_slot))

    def dt(self, acm_c):
        out = ctypes.c_double()
        check(load().wh_dt_reduce(self.h, ctypes.byref(acm_c), ctypes.byref(out)))
        return out.value

    def divergence_guard(self, array=WH_BLOCK):
        out = ctypes.c_double()
        check(load().wh_divergence_guard(self.h, array, ctypes.byref(out)))
        return out.value

    def set_block_groups(self, groups, ngroups):
        """Group (independent forest) of every active block, in active-list order."""
        g = np.ascontiguousarray(groups, np.int32)
        self.ngroups = int(ngroups)
        check(load().wh_set_block_groups(self.h, g.ctypes.data if g.size else None, g.size, int(ngroups)))

    def dt_groups(self, acm_c):
        """Per-group minimum block dt (before the host clamps)."""
        out = np.zeros(self.ngroups)
        check(load().wh_dt_reduce_groups(self.h, ctypes.byref(acm_c), out.ctypes.data))
        return out

    def rk_step_groups(self, acm_c, butcher, dts, path=2):
        bt = np.ascontiguousarray(butcher, np.float64)
        d = np.ascontiguousarray(dts, np.float64)
        check(load().wh_rk_step_groups(self.h, ctypes.byref(acm_c), bt.ctypes.data, bt.shape[0],
                                       d.ctypes.data, path))

    def set_group_params(self, params):
        """Physics constants per block group (a list of AcmParamsC, one per
        group); an empty list clears them."""
        arr = (AcmParamsC * max(1, len(params)))(*params) if params else None
        check(load().wh_set_group_params(self.h, ctypes.cast(arr, ctypes.c_void_p) if arr is not None else None,
                                         len(params)))

    def rk_step(self, acm_c, butcher, time, dt, path=0):
        bt = np.ascontiguousarray(butcher, np.float64)
        check(load().wh_rk_step(self.h, ctypes.byref(acm_c), bt.ctypes.data, bt.shape[0], time, dt, path))

    def time_set(self, time, iteration):
        check(load().wh_time_set(self.h, time, iteration))

    def time_get(self):
        t, it, dt = ctypes.c_double(), ctypes.c_int(), ctypes.c_double()
        check(load().wh_time_get(self.h, ctypes.byref(t), ctypes.byref(it), ctypes.byref(dt)))
        return t.value, it.value, dt.value

    def time_stage_kernel(self, acm_c, reps=20):
        ms = ctypes.c_double()
        check(load().wh_time_stage_kernel(self.h, ctypes.byref(acm_c), reps, ctypes.byref(ms)))
        return ms.value

    def time_fetch(self):
        check(load().wh_time_fetch(self.h))

    def time_fetched(self):
        t, it, dt = ctypes.c_double(), ctypes.c_int(), ctypes.c_double()
        check(load().wh_time_fetched(self.h, ctypes.byref(t), ctypes.byref(it), ctypes.byref(dt)))
        return t.value, it.value, dt.value

    def run_steps(self, acm_c, butcher, tp_c, nsteps, use_graph=True):
        bt = np.ascontiguousarray(butcher, np.float64)
        check(load().wh_run_steps(self.h, ctypes.byref(acm_c), bt.ctypes.data, bt.shape[0],
                                  ctypes.byref(tp_c), nsteps, int(use_graph)))

    def synchronize(self):
        check(load().wh_synchronize(self.h))

    def set_tuning(self, key, value):
        check(load().wh_set_tuning(self.h, key, value))

    def get_tuning(self, key):
        v = ctypes.c_int()
        check(load().wh_get_tuning(self.h, key, ctypes.byref(v)))
        return v.value

    def autotune(self, acm_c, reps=5):
        v, ms = ctypes.c_int(), ctypes.c_double()
        check(load().wh_autotune(self.h, ctypes.byref(acm_c), reps, ctypes.byref(v), ctypes.byref(ms)))
        return v.value, ms.value

    def timing(self, on):
        check(load().wh_kernel_timing_enable(self.h, int(on)))
        self.timing_on = bool(on)
        check(load().wh_kernel_timing(self.h, -1, None, None))

    def kernel_time(self, which):
        ms, n = ctypes.c_double(), ctypes.c_int()
        check(load().wh_kernel_timing(self.h, which, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


class BatchStepper:
    """RungeKuttaGeneric over the block groups of several pools of one
    dimension in one launch per RK stage (wh_batch_*): each pool's active
    list, patch table and block groups must be set before each step."""

    def __init__(self, pools):
        self.pools = list(pools)
        arr = (ctypes.c_void_p * len(self.pools))(*[p.h.value for p in self.pools])
        h = ctypes.c_void_p()
        check(load().wh_batch_create(ctypes.byref(h), ctypes.cast(arr, ctypes.c_void_p), len(self.pools)))
        self.h = h

    def rk_step(self, acm_c, butcher, dt_groups, use_graph=True):
        bt = np.ascontiguousarray(butcher, np.float64)
        d = np.ascontiguousarray(dt_groups, np.float64)
        check(load().wh_batch_rk_step(self.h, ctypes.byref(acm_c), bt.ctypes.data, bt.shape[0],
                                      d.ctypes.data if d.size else None, int(use_graph)))

    def time_stage_kernel(self, acm_c, butcher, reps=5):
        """(ms per launch, blocks per launch) of the fused stage kernel."""
        bt = np.ascontiguousarray(butcher, np.float64)
        ms, nb = ctypes.c_double(), ctypes.c_int()
        check(load().wh_batch_time_stage_kernel(self.h, ctypes.byref(acm_c), bt.ctypes.data, bt.shape[0], reps,
                                                ctypes.byref(ms), ctypes.byref(nb)))
        return ms.value, nb.value

    def synchronize(self):
        check(load().wh_batch_synchronize(self.h))

    def set_tuning(self, key, value):
        check(load().wh_batch_set_tuning(self.h, key, value))

    def get_tuning(self, key):
        v = ctypes.c_int()
        check(load().wh_batch_get_tuning(self.h, key, ctypes.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "h", None):
            load().wh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
