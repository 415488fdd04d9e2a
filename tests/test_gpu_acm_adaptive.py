"""GPU: the adaptive penalised-cylinder run (the reference's
TESTING/acm/acm_CDF44, see tests/acm_cyl_case.py) on the device, against the
oracle's run and the reference's saved states at t = 0.05 and 0.1.

Bars: block lists (treecodes, levels) and iteration counts identical; fields
within relative L2 1e-12 of both (the device's cos/pow in the mask may differ
from the host libm by an ulp, so the fields are not compared bit for bit)."""
import numpy as np
import pytest

import acm_cyl_case as A
from helpers import load_npz
from oracle import pyoracle as po
from wabbit_amd import lib as L
from wabbit_amd import wavelets as W
from wabbit_amd.adaptive import AdaptConfig, AdaptiveACM

pytestmark = pytest.mark.gpu

G, BS = A.G, A.BS


def device_mask_params():
    w, s = A.smoothing()
    return W.mask_params(geometry=1, x_cntr=(10.0, 10.0, 0.0), R_cyl=0.5, smoothing_width=w, smoothing_safety=s,
                         use_sponge=True, sponge_type=1, L_sponge=2.0, p_sponge=8.0, domain=(A.L, A.L, 1.0))


def solver(path=2, decompose_branch=0):
    tc, lv, hvy, jmax_tc = A.initial_state()
    cfg = AdaptConfig(wavelet="CDF44", bs=BS, jmin=A.JMIN, jmax=A.JMAX, eps=1e-3, eps_normalized=True,
                      groups=[1, 1, 1], threshold_mask=True, force_maxlevel_dealiasing=True,
                      decompose_branch=decompose_branch)
    return AdaptiveACM(A.acm_params(), A.time_params(), cfg, tc, lv, hvy, mask=device_mask_params(),
                       nb_max=2500, path=path, jmax_treecode=jmax_tc)


def test_mask_kernel_matches_oracle():
    s = solver()
    tc, lv, _ = s.state()
    got = s.mask()
    hvy = np.zeros((len(tc), 3, 1, BS + 2 * G, BS + 2 * G))
    tree = A.oracle_tree(tc, lv, hvy, A.JMAX)
    want = tree.get_mask(A.oracle_mask_params())
    assert got.shape == want.shape
    # chi and colour within an ulp; the p-norm sponge step_cos(tmp) with
    # tmp = -(pow(pow(x,8)+pow(y,8),1/8) - L/2): the device and host pow may
    # differ by an ulp or two of |tmp| <= L/2 = 10 (ulp 1.8e-15), scaled by
    # the cosine step's slope pi/(2h) = pi/2 for h = L_sponge/2 = 1
    d = np.abs(got - want)
    assert np.max(d[:, [0, 1, 2, 3, 4]]) <= 4e-16
    assert np.max(d[:, 5]) <= 1e-14
    # and chi on the reference's saved grid
    ref = load_npz("acm_cyl_t0.npz")
    key = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(tc, lv))}
    chi = np.stack([got[key[(int(a), int(b))], 0, 0, G:G + BS + 1, G:G + BS + 1]
                    for a, b in zip(ref["treecode"], ref["level"])])
    assert np.max(np.abs(chi - ref["field_mask"])) <= 4e-16
    s.close()


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _check(s, ref, tree=None):
    tc, lv, data = s.state()
    keys = [(int(a), int(b)) for a, b in zip(tc, lv)]
    assert sorted(keys) == sorted((int(a), int(b)) for a, b in zip(ref["treecode"], ref["level"]))
    key = {k: i for i, k in enumerate(keys)}
    for c, name in enumerate(["ux", "uy", "p"]):
        mine = np.stack([data[key[(int(a), int(b))], c, 0, G:G + BS, G:G + BS]
                         for a, b in zip(ref["treecode"], ref["level"])])
        assert _rel(mine, ref["field_" + name][:, :BS, :BS]) <= 1e-12, name
    if tree is not None:
        tco, lvo, hvo = tree.get()
        assert np.array_equal(tco, tc) and np.array_equal(lvo, lv)
        inner = (slice(None), slice(None), 0, slice(G, G + BS), slice(G, G + BS))
        assert _rel(data[inner], hvo[inner]) <= 1e-12


@pytest.mark.parametrize("path", [2, 0])
def test_adaptive_cylinder_run(path):
    s = solver(path)
    ref05, ref10 = load_npz("acm_cyl_t005.npz"), load_npz("acm_cyl_t010.npz")
    oracle = {}

    def grab(tree, time, it):
        if abs(time - 0.05) < 1e-12:
            oracle["t005"] = tree.get()
    tree, _, it_o = A.oracle_run(0.1, grab)
    while s.time < 0.05 - 1e-12:
        s.step()
    assert s.iteration == int(ref05["iteration"])
    _check(s, ref05)
    tc, lv, data = s.state()
    tco, lvo, hvo = oracle["t005"]
    assert np.array_equal(tco, tc) and np.array_equal(lvo, lv)
    while s.time < 0.1 - 1e-12:
        s.step()
    assert s.iteration == int(ref10["iteration"]) == it_o
    _check(s, ref10, tree)
    s.close()


def test_level_wise_branch_run():
    """The level-wise decomposition forced on the device (Bs = 26 takes the
    leaf-first branch): the reference's saved state at t = 0.05."""
    s = solver(2, decompose_branch=2)
    ref05 = load_npz("acm_cyl_t005.npz")
    while s.time < 0.05 - 1e-12:
        s.step()
    assert s.iteration == int(ref05["iteration"])
    _check(s, ref05)
    s.close()


def test_save_and_resume(tmp_path):
    """save_data at t = 0.05 (one WABBIT HDF5 file per component), a fresh
    solver resumed from the files (read_from_files), on to t = 0.1: the
    same result as the uninterrupted run, bit for bit, and the reference's
    saved state."""
    from wabbit_amd import h5io
    if not h5io.available():
        pytest.skip("libhdf5 not found")
    a = solver(2)
    while a.time < 0.05 - 1e-12:
        a.step()
    paths = a.save(str(tmp_path))
    assert [p.rsplit("/", 1)[1] for p in paths] == ["ux_000000050000.h5", "uy_000000050000.h5",
                                                     "p_000000050000.h5"]
    f = h5io.read_wabbit_file(paths[0])
    ref05 = load_npz("acm_cyl_t005.npz")
    key = {(int(x), int(y)): k for k, (x, y) in enumerate(zip(f.treecode, f.level))}
    mine = np.stack([f.blocks[key[(int(x), int(y))]] for x, y in zip(ref05["treecode"], ref05["level"])])
    assert _rel(mine[:, :BS, :BS], ref05["field_ux"][:, :BS, :BS]) <= 1e-12
    b = solver(2)
    b.load(paths)
    assert (b.time, b.iteration) == (a.time, a.iteration)
    while a.time < 0.1 - 1e-12:
        a.step()
        b.step()
    assert (b.time, b.iteration) == (a.time, a.iteration)
    ta, la, da = a.state()
    tb, lb, db = b.state()
    assert np.array_equal(ta, tb) and np.array_equal(la, lb)
    inner = (slice(None), slice(None), 0, slice(G, G + BS), slice(G, G + BS))
    assert np.array_equal(da[inner], db[inner])
    _check(b, load_npz("acm_cyl_t010.npz"))
    a.close()
    b.close()
