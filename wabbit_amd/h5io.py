"""Reader/writer for WABBIT's HDF5 block files, via ctypes on libhdf5.

One file holds one field of a forest (LIB/MESH/InputOutput.f90:1-309,
saveHDF5_tree; :322-764, readHDF5vct_tree):

  blocks              f64  (N, Bs2+1, Bs1+1)      2D, or (N, Bs3+1, Bs2+1, Bs1+1) 3D:
                           the interior plus the first upper ghost point
                           (InputOutput.f90:174,203)
  block_treecode_num  i64  (N,)
  level, refinement_status, procs, lgt_ids  i32 (N,)
  coords_origin, coords_spacing  f64 (N, dim), stored in reversed axis order
                           (z, y, x) -- InputOutput.f90:207-211
  attributes on "blocks": time, iteration, dim, domain-size, block-size,
                           max_level, periodic_BC, symmetry_BC,
                           total_number_blocks, version

The arrays come back in C order, so ``blocks[n, iy, ix]`` is the Fortran
``hvy_block(g+1+ix, g+1+iy, 1, dF, n)``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_hid_t = ctypes.c_int64
_herr_t = ctypes.c_int
_hsize_t = ctypes.c_uint64

_LIB = None


def _lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    cands = [os.environ.get("WABBIT_HDF5_LIB", ""), "/opt/conda/lib/libhdf5.so.103",
             "/opt/conda/lib/libhdf5.so", "libhdf5.so"]
    err = None
    for c in cands:
        if not c:
            continue
        try:
            lib = ctypes.CDLL(c)
            break
        except OSError as e:  # pragma: no cover - depends on the image
            err = e
    else:
        raise OSError(f"libhdf5 not found ({err})")
    lib.H5open.restype = _herr_t
    lib.H5open()
    # optional datasets/attributes are probed; keep HDF5's error stack quiet
    lib.H5Eset_auto2.argtypes = [_hid_t, ctypes.c_void_p, ctypes.c_void_p]
    lib.H5Eset_auto2(0, None, None)
    lib.H5Fopen.restype = _hid_t
    lib.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, _hid_t]
    lib.H5Fcreate.restype = _hid_t
    lib.H5Fcreate.argtypes = [ctypes.c_char_p, ctypes.c_uint, _hid_t, _hid_t]
    lib.H5Fclose.argtypes = [_hid_t]
    lib.H5Dopen2.restype = _hid_t
    lib.H5Dopen2.argtypes = [_hid_t, ctypes.c_char_p, _hid_t]
    lib.H5Dcreate2.restype = _hid_t
    lib.H5Dcreate2.argtypes = [_hid_t, ctypes.c_char_p, _hid_t, _hid_t, _hid_t, _hid_t, _hid_t]
    lib.H5Dclose.argtypes = [_hid_t]
    lib.H5Dget_space.restype = _hid_t
    lib.H5Dget_space.argtypes = [_hid_t]
    lib.H5Dread.argtypes = [_hid_t, _hid_t, _hid_t, _hid_t, _hid_t, ctypes.c_void_p]
    lib.H5Dwrite.argtypes = [_hid_t, _hid_t, _hid_t, _hid_t, _hid_t, ctypes.c_void_p]
    lib.H5Sget_simple_extent_ndims.argtypes = [_hid_t]
    lib.H5Sget_simple_extent_dims.argtypes = [_hid_t, ctypes.POINTER(_hsize_t), ctypes.POINTER(_hsize_t)]
    lib.H5Screate_simple.restype = _hid_t
    lib.H5Screate_simple.argtypes = [ctypes.c_int, ctypes.POINTER(_hsize_t), ctypes.POINTER(_hsize_t)]
    lib.H5Sclose.argtypes = [_hid_t]
    lib.H5Aopen_by_name.restype = _hid_t
    lib.H5Aopen_by_name.argtypes = [_hid_t, ctypes.c_char_p, ctypes.c_char_p, _hid_t, _hid_t]
    lib.H5Acreate2.restype = _hid_t
    lib.H5Acreate2.argtypes = [_hid_t, ctypes.c_char_p, _hid_t, _hid_t, _hid_t, _hid_t]
    lib.H5Aget_space.restype = _hid_t
    lib.H5Aget_space.argtypes = [_hid_t]
    lib.H5Aread.argtypes = [_hid_t, _hid_t, ctypes.c_void_p]
    lib.H5Awrite.argtypes = [_hid_t, _hid_t, ctypes.c_void_p]
    lib.H5Aclose.argtypes = [_hid_t]
    _LIB = lib
    return lib


def available() -> bool:
    """True if a libhdf5 shared library can be loaded."""
    try:
        _lib()
        return True
    except OSError:
        return False


def _type(name):
    return _hid_t.in_dll(_lib(), name).value


_NATIVE = {
    np.dtype(np.float64): "H5T_NATIVE_DOUBLE_g",
    np.dtype(np.int32): "H5T_NATIVE_INT_g",
    np.dtype(np.int64): "H5T_NATIVE_LLONG_g",
}
_FILE_T = {
    np.dtype(np.float64): "H5T_IEEE_F64LE_g",
    np.dtype(np.int32): "H5T_STD_I32LE_g",
    np.dtype(np.int64): "H5T_STD_I64LE_g",
}


def _extent(space):
    lib = _lib()
    nd = lib.H5Sget_simple_extent_ndims(space)
    dims = (_hsize_t * max(nd, 1))()
    lib.H5Sget_simple_extent_dims(space, dims, None)
    return tuple(int(dims[i]) for i in range(nd))


def _read_dset(fid, name, dtype):
    lib = _lib()
    did = lib.H5Dopen2(fid, name.encode(), 0)
    if did < 0:
        raise KeyError(name)
    sp = lib.H5Dget_space(did)
    shape = _extent(sp)
    out = np.empty(shape, dtype=dtype)
    if lib.H5Dread(did, _type(_NATIVE[np.dtype(dtype)]), 0, 0, 0, out.ctypes.data) < 0:
        raise IOError(f"H5Dread {name}")
    lib.H5Sclose(sp)
    lib.H5Dclose(did)
    return out


def _read_attr(fid, obj, name, dtype):
    lib = _lib()
    aid = lib.H5Aopen_by_name(fid, obj.encode(), name.encode(), 0, 0)
    if aid < 0:
        raise KeyError(name)
    sp = lib.H5Aget_space(aid)
    shape = _extent(sp) or (1,)
    out = np.empty(shape, dtype=dtype)
    lib.H5Aread(aid, _type(_NATIVE[np.dtype(dtype)]), out.ctypes.data)
    lib.H5Sclose(sp)
    lib.H5Aclose(aid)
    return out


@dataclass
class WabbitFile:
    """Contents of one WABBIT field file (see module docstring)."""
    blocks: np.ndarray
    treecode: np.ndarray
    level: np.ndarray
    refinement_status: np.ndarray
    origin: np.ndarray      # (N, dim) in (x, y[, z]) order
    spacing: np.ndarray     # (N, dim) in (x, y[, z]) order
    time: float
    iteration: int
    dim: int
    domain_size: np.ndarray
    bs: np.ndarray
    max_level: int
    periodic_bc: np.ndarray = field(default_factory=lambda: np.ones(3, np.int32))
    version: int = 0


def read_wabbit_file(path: str) -> WabbitFile:
    lib = _lib()
    fid = lib.H5Fopen(path.encode(), 0, 0)
    if fid < 0:
        raise IOError(f"cannot open {path}")
    try:
        blocks = _read_dset(fid, "blocks", np.float64)
        dim = blocks.ndim - 1
        try:
            dim = int(_read_attr(fid, "blocks", "dim", np.int32)[0])
        except KeyError:
            pass
        try:
            tc = _read_dset(fid, "block_treecode_num", np.int64)
            lev = _read_dset(fid, "level", np.int32)
            ml = int(_read_attr(fid, "blocks", "max_level", np.int32)[0])
        except KeyError:
            # files before version 20240410 store the treecode as a digit array
            # (N, tc_length), -1 past the block's level (InputOutput.f90:491-558,
            # treecode_size / array2tcb)
            digits = _read_dset(fid, "block_treecode", np.int32)
            ml = digits.shape[1]
            lev = (digits >= 0).sum(axis=1).astype(np.int32)
            tc = np.zeros(digits.shape[0], np.int64)
            for l in range(ml):
                d = digits[:, l].astype(np.int64)
                tc |= np.where(d >= 0, d, 0) << ((ml - 1 - l) * dim)
        rs = _read_dset(fid, "refinement_status", np.int32)
        org = _read_dset(fid, "coords_origin", np.float64)[:, ::-1].copy()
        spc = _read_dset(fid, "coords_spacing", np.float64)[:, ::-1].copy()
        time = float(_read_attr(fid, "blocks", "time", np.float64)[0])
        it = int(_read_attr(fid, "blocks", "iteration", np.int32)[0])
        dom = _read_attr(fid, "blocks", "domain-size", np.float64)
        bs = _read_attr(fid, "blocks", "block-size", np.int32)
        try:
            per = _read_attr(fid, "blocks", "periodic_BC", np.int32)
        except KeyError:
            per = np.ones(3, np.int32)
        try:
            ver = int(_read_attr(fid, "blocks", "version", np.int32)[0])
        except KeyError:
            ver = 0
    finally:
        lib.H5Fclose(fid)
    return WabbitFile(blocks, tc, lev, rs, org, spc, time, it, dim, dom, bs, ml, per, ver)


def _write_dset(fid, name, arr):
    lib = _lib()
    arr = np.ascontiguousarray(arr)
    dims = (_hsize_t * arr.ndim)(*arr.shape)
    sp = lib.H5Screate_simple(arr.ndim, dims, None)
    did = lib.H5Dcreate2(fid, name.encode(), _type(_FILE_T[arr.dtype]), sp, 0, 0, 0)
    lib.H5Dwrite(did, _type(_NATIVE[arr.dtype]), 0, 0, 0, arr.ctypes.data)
    lib.H5Dclose(did)
    lib.H5Sclose(sp)


def _write_attr(fid, obj, name, arr):
    lib = _lib()
    arr = np.ascontiguousarray(np.atleast_1d(arr))
    dims = (_hsize_t * 1)(arr.size)
    sp = lib.H5Screate_simple(1, dims, None)
    did = lib.H5Dopen2(fid, obj.encode(), 0)
    aid = lib.H5Acreate2(did, name.encode(), _type(_FILE_T[arr.dtype]), sp, 0, 0)
    lib.H5Awrite(aid, _type(_NATIVE[arr.dtype]), arr.ctypes.data)
    lib.H5Aclose(aid)
    lib.H5Dclose(did)
    lib.H5Sclose(sp)


def write_wabbit_file(path: str, f: WabbitFile) -> None:
    """Write ``f`` in the layout saveHDF5_tree produces (InputOutput.f90:233-283)."""
    lib = _lib()
    fid = lib.H5Fcreate(path.encode(), 2, 0, 0)  # H5F_ACC_TRUNC
    if fid < 0:
        raise IOError(f"cannot create {path}")
    try:
        _write_dset(fid, "blocks", f.blocks.astype(np.float64))
        _write_dset(fid, "coords_origin", np.ascontiguousarray(f.origin[:, ::-1], dtype=np.float64))
        _write_dset(fid, "coords_spacing", np.ascontiguousarray(f.spacing[:, ::-1], dtype=np.float64))
        _write_dset(fid, "block_treecode_num", f.treecode.astype(np.int64))
        n = f.blocks.shape[0]
        _write_dset(fid, "procs", np.zeros(n, np.int32))
        _write_dset(fid, "refinement_status", f.refinement_status.astype(np.int32))
        _write_dset(fid, "level", f.level.astype(np.int32))
        _write_dset(fid, "lgt_ids", np.arange(1, n + 1, dtype=np.int32))
        _write_attr(fid, "blocks", "domain-size", np.asarray(f.domain_size, np.float64)[: f.dim])
        _write_attr(fid, "blocks", "periodic_BC", np.asarray(f.periodic_bc, np.int32))
        _write_attr(fid, "blocks", "symmetry_BC", np.zeros(3, np.int32))
        _write_attr(fid, "blocks", "version", np.array([20240410], np.int32))
        _write_attr(fid, "blocks", "block-size", np.asarray(f.bs, np.int32))
        _write_attr(fid, "blocks", "time", np.array([f.time], np.float64))
        _write_attr(fid, "blocks", "iteration", np.array([f.iteration], np.int32))
        _write_attr(fid, "blocks", "total_number_blocks", np.array([n], np.int32))
        _write_attr(fid, "blocks", "max_level", np.array([f.max_level], np.int32))
        _write_attr(fid, "blocks", "dim", np.array([f.dim], np.int32))
    finally:
        lib.H5Fclose(fid)


def files_from_state(treecodes, levels, data, g, bs, dim, domain, max_level, time, iteration):
    """One WabbitFile per component of a leaf state (data: (n, neq, nz, ny,
    nx) with synchronised ghost nodes) in saveHDF5_tree's layout: each block
    stores its interior plus the first upper ghost point (InputOutput.f90:
    174-211).  Origins and spacings follow get_block_spacing_origin."""
    from .forest import ijk_from_treecode, spacing_origin
    tc = np.asarray(treecodes, np.int64)
    lv = np.asarray(levels, np.int32)
    n = len(tc)
    bsl = [int(b) for b in (list(bs) + [1] * 3)[:3]]
    org = np.zeros((n, dim))
    spc = np.zeros((n, dim))
    for k in range(n):
        ijk = ijk_from_treecode(int(tc[k]), int(lv[k]), max_level, dim)
        x0, dx = spacing_origin(ijk, int(lv[k]), domain, bsl[:dim], dim)
        org[k], spc[k] = x0[:dim], dx[:dim]
    if dim == 2:
        sl = (slice(None), 0, slice(g, g + bsl[1] + 1), slice(g, g + bsl[0] + 1))
    else:
        sl = (slice(None), slice(g, g + bsl[2] + 1), slice(g, g + bsl[1] + 1), slice(g, g + bsl[0] + 1))
    out = []
    for c in range(data.shape[1]):
        blocks = np.ascontiguousarray(data[:, c][sl])
        out.append(WabbitFile(blocks=blocks, treecode=tc.copy(), level=lv.copy(),
                              refinement_status=np.zeros(n, np.int32), origin=org, spacing=spc,
                              time=float(time), iteration=int(iteration), dim=dim,
                              domain_size=np.asarray(list(domain)[:dim], np.float64),
                              bs=np.asarray(bsl[:dim], np.int32), max_level=int(max_level)))
    return out


def state_from_files(files, g, max_level=None):
    """readHDF5vct_tree (InputOutput.f90:322-764) into a leaf state: (treecodes,
    levels, data (n, neq, nz, ny, nx), time, iteration).  Each file is one
    component; the stored interior (plus first upper ghost point) goes to
    the block's interior, every other ghost node is left 0 for the caller's
    sync_ghosts_tree.  Treecodes are re-based to ``max_level`` digits if
    given (the digit of level l sits at bit (max_level - l)*dim)."""
    f0 = files[0]
    for f in files[1:]:
        if not (np.array_equal(f.treecode, f0.treecode) and np.array_equal(f.level, f0.level)):
            raise ValueError("component files hold different grids")
    dim = f0.dim
    bsl = [int(b) for b in f0.bs] + [1] * (3 - dim)
    n = len(f0.treecode)
    nx, ny = bsl[0] + 2 * g, bsl[1] + 2 * g
    nz = bsl[2] + 2 * g if dim == 3 else 1
    data = np.zeros((n, len(files), nz, ny, nx))
    for c, f in enumerate(files):
        if dim == 2:
            data[:, c, 0, g:g + bsl[1] + 1, g:g + bsl[0] + 1] = f.blocks
        else:
            data[:, c, g:g + bsl[2] + 1, g:g + bsl[1] + 1, g:g + bsl[0] + 1] = f.blocks
    tc = f0.treecode.astype(np.int64)
    if max_level is not None and max_level != f0.max_level:
        sh = (int(max_level) - int(f0.max_level)) * dim
        tc = tc << sh if sh >= 0 else tc >> (-sh)
    return tc, f0.level.astype(np.int32), data, f0.time, f0.iteration


def timestr(time: float) -> str:
    """The 12-digit time stamp of the reference's file names (module_helpers.f90
    timestr: whole part i6.6, then the fraction in millionths i6.6)."""
    import math
    whole = int(time + 1e-12)
    frac = int(round(max((time - math.floor(time + 1e-12)) * 1e6, 0.0)))
    return f"{whole:06d}{frac:06d}"
