"""WABBIT's HDF5 field files (saveHDF5_tree / readHDF5vct_tree,
InputOutput.f90:1-764): the writer against the reader, the reference's own
saved states through both, and the leaf-state <-> file conversion the
save/resume path uses."""
import os

import numpy as np
import pytest

from helpers import load_npz
from wabbit_amd import h5io
from wabbit_amd.forest import ijk_from_treecode, spacing_origin

pytestmark = pytest.mark.skipif(not h5io.available(), reason="libhdf5 not found")


def _fixture_file(name, field):
    f = load_npz(name)
    dim = int(f["dim"]) if "dim" in f else 2
    return h5io.WabbitFile(blocks=np.ascontiguousarray(f["field_" + field]), treecode=f["treecode"].astype(np.int64),
                           level=f["level"].astype(np.int32), refinement_status=np.zeros(len(f["level"]), np.int32),
                           origin=np.ascontiguousarray(f["origin"]), spacing=np.ascontiguousarray(f["spacing"]),
                           time=float(f["time"]), iteration=int(f["iteration"]), dim=dim,
                           domain_size=np.asarray(f["domain_size"], np.float64), bs=np.asarray(f["bs"], np.int32),
                           max_level=int(f["max_level"]))


def _same(a, b):
    for k in ("blocks", "treecode", "level", "refinement_status", "origin", "spacing", "domain_size", "bs"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert a.time == b.time and a.iteration == b.iteration and a.dim == b.dim and a.max_level == b.max_level


@pytest.mark.parametrize("name,field", [("acm_cyl_t005.npz", "ux"), ("acm_cyl_t010.npz", "p"),
                                        ("wavelet_input.npz", "vor")])
def test_reference_state_round_trip(tmp_path, name, field):
    """A state the reference saved (converted from its .h5 files) written
    and read back is unchanged, bit for bit."""
    f = _fixture_file(name, field)
    path = os.path.join(tmp_path, "f.h5")
    h5io.write_wabbit_file(path, f)
    g = h5io.read_wabbit_file(path)
    _same(f, g)
    assert g.version == 20240410


@pytest.mark.parametrize("dim", [2, 3])
def test_state_files_state(tmp_path, dim):
    """leaf state -> saveHDF5_tree layout -> readHDF5vct_tree -> leaf state:
    interiors and the first upper ghost point come back exactly, the other
    ghost nodes are zero; origins/spacings are get_block_spacing_origin's."""
    rng = np.random.default_rng(3)
    jmax, bs, g = 3, 8, 3
    tc = np.array([0, 1 << dim, 5 << dim, (1 << (2 * dim)) - 1], np.int64) if dim == 2 else \
        np.array([0, 9, 17 << 3, 63], np.int64)
    lv = np.array([2, 3, 3, 3], np.int32) if dim == 2 else np.array([3, 3, 2, 3], np.int32)
    tc = tc << 0
    n1 = bs + 2 * g
    shape = (len(tc), dim + 1, n1 if dim == 3 else 1, n1, n1)
    data = rng.standard_normal(shape)
    dom = [4.0, 2.0, 1.0]
    files = h5io.files_from_state(tc, lv, data, g, [bs] * dim, dim, dom, jmax, 0.125, 17)
    paths = []
    for c, f in enumerate(files):
        p = os.path.join(tmp_path, f"c{c}_{h5io.timestr(0.125)}.h5")
        h5io.write_wabbit_file(p, f)
        paths.append(p)
    assert os.path.basename(paths[0]) == "c0_000000125000.h5"
    back = [h5io.read_wabbit_file(p) for p in paths]
    tc2, lv2, d2, t2, it2 = h5io.state_from_files(back, g)
    assert np.array_equal(tc2, tc) and np.array_equal(lv2, lv) and (t2, it2) == (0.125, 17)
    keep = np.zeros(shape[2:], bool)
    if dim == 2:
        keep[0, g:g + bs + 1, g:g + bs + 1] = True
    else:
        keep[g:g + bs + 1, g:g + bs + 1, g:g + bs + 1] = True
    assert np.array_equal(d2[:, :, keep], data[:, :, keep])
    assert not np.any(d2[:, :, ~keep])
    for k in range(len(tc)):
        x0, dx = spacing_origin(ijk_from_treecode(int(tc[k]), int(lv[k]), jmax, dim), int(lv[k]), dom, [bs] * dim,
                                dim)
        assert np.array_equal(back[0].origin[k], x0[:dim]) and np.array_equal(back[0].spacing[k], dx[:dim])


def test_treecode_rebased_to_other_max_level():
    """A file written with max_level 3 read into a forest of max_level 5:
    every digit moves up by (5 - 3) * dim bits."""
    f = _fixture_file("acm_cyl_t005.npz", "ux")
    tc, lv, _, _, _ = h5io.state_from_files([f], 6, max_level=int(f.max_level) + 2)
    assert np.array_equal(tc, f.treecode << 4)


def test_timestr():
    assert h5io.timestr(0.0) == "000000000000"
    assert h5io.timestr(10.05) == "000010050000"
    assert h5io.timestr(123.4567891) == "000123456789"
