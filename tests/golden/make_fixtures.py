"""Convert the reference's own regression data (HDF5) into small .npz fixtures.

Runs only where the reference tree is mounted (read-only, at
/root/reference); the GPU machine only sees the .npz files this writes.  The
data are the reference's committed golden inputs/outputs, unchanged:

  vortex2d_ic.npz          TESTING/acm/3vortices/{ux,uy,p}_000010000000.h5
                           (t=10, iteration 3054; shared input of the three
                           3vorticesEqui* tests)
  vortex2d_fd{2,4,6}.npz   TESTING/acm/3vortices/3vorticesEquiFD{2,4,6}_CDF{20,40,60}/
                           {ux,uy,p}_000020000000.h5 (t=20, iteration 6127)
  tg3d_fd{2,4,6}.npz       TESTING/acm/taylorGreen/taylorGreenEqui_FD{2,4,6}_CDF{20,40,60}/
                           {ux,uy,uz,p}_00000{0,10}000000.h5 (t=0 input, t=10 output)

Each array keeps the file's (N, Bs+1, ..) storage: the interior plus the
first upper ghost point.  Parameters of each run are copied from its
PARAMS_*.ini into the fixture's ``params_json`` entry.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from wabbit_amd.h5io import read_wabbit_file  # noqa: E402

REF = "/root/reference/TESTING/acm"


def _fields(dirpath, names, tag):
    out = {}
    meta = None
    for n in names:
        f = read_wabbit_file(os.path.join(dirpath, f"{n}_{tag}.h5"))
        out[n] = f.blocks
        if meta is None:
            meta = f
        else:
            assert np.array_equal(meta.treecode, f.treecode)
    return out, meta


def _save(path, fields, meta, params):
    np.savez_compressed(
        path,
        **{f"field_{k}": v for k, v in fields.items()},
        treecode=meta.treecode, level=meta.level, origin=meta.origin,
        spacing=meta.spacing, time=meta.time, iteration=meta.iteration,
        dim=meta.dim, domain_size=meta.domain_size, bs=meta.bs,
        max_level=meta.max_level,
        params_json=json.dumps(params),
    )
    print("wrote", path, os.path.getsize(path))


def vortex2d():
    names = ["ux", "uy", "p"]
    ic, meta = _fields(os.path.join(REF, "3vortices"), names, "000010000000")
    _save(os.path.join(HERE, "vortex2d_ic.npz"), ic, meta, {})
    for order, wav in ((2, "CDF20"), (4, "CDF40"), (6, "CDF60")):
        d = os.path.join(REF, "3vortices", f"3vorticesEquiFD{order}_{wav}")
        # all three runs start from the shared input
        ic2, _ = _fields(d, names, "000010000000")
        for n in names:
            assert np.array_equal(ic2[n], ic[n])
        out, m = _fields(d, names, "000020000000")
        assert np.array_equal(m.treecode, meta.treecode)
        params = dict(scheme=order, wavelet=wav, c_0=7.0, nu=5.0e-5, gamma_p=1.0,
                      CFL=1.0, skew_symmetry=1, write_time=10.0, time_max=20.0,
                      Bs=32, Jmax=3, L=float(meta.domain_size[0]))
        _save(os.path.join(HERE, f"vortex2d_fd{order}.npz"), out, m, params)


def taylor_green():
    names = ["ux", "uy", "uz", "p"]
    for order, wav in ((2, "CDF20"), (4, "CDF40"), (6, "CDF60")):
        d = os.path.join(REF, "taylorGreen", f"taylorGreenEqui_FD{order}_{wav}")
        ic, meta = _fields(d, names, "000000000000")
        out, m = _fields(d, names, "000010000000")
        assert np.array_equal(m.treecode, meta.treecode)
        params = dict(scheme=order, wavelet=wav, c_0=10.0, nu=3.125e-3, gamma_p=0.0,
                      CFL=1.0, skew_symmetry=1, write_time=10.0, time_max=10.0,
                      tsave_stats=0.2, Bs=20, Jmax=1, L=float(meta.domain_size[0]),
                      t0=meta.time, it0=int(meta.iteration), t1=m.time, it1=int(m.iteration))
        fields = {f"ic_{k}": v for k, v in ic.items()}
        fields.update({f"out_{k}": v for k, v in out.items()})
        _save(os.path.join(HERE, f"tg3d_fd{order}.npz"), fields, m, params)


def wavelets():
    """TESTING/wavelets: input vor_000020000000.h5 (112 blocks, levels 2-5,
    Bs=32, old digit-array treecodes), the --refine-everywhere outputs
    (vor_00100.h5, kept by the reference for CDF20/40/60; CDF20 and CDF60
    subsampled to every 8th block) and the --coarsen-everywhere outputs of the
    lifted wavelets (vor_00200.h5; for CDF20/40/60 it equals the input bit for
    bit, asserted here)."""
    W = "/root/reference/TESTING/wavelets"
    src = read_wabbit_file(os.path.join(W, "vor_000020000000.h5"))
    _save(os.path.join(HERE, "wavelet_input.npz"), {"vor": src.blocks}, src, {})
    for wav in ("CDF20", "CDF40", "CDF60"):
        f = read_wabbit_file(os.path.join(W, f"adaptive_{wav}", "vor_00100.h5"))
        keep = np.arange(f.blocks.shape[0])
        if wav != "CDF40":
            keep = keep[::8]
        sub = type(f)(**{**f.__dict__})
        sub.blocks, sub.treecode, sub.level = f.blocks[keep], f.treecode[keep], f.level[keep]
        sub.origin, sub.spacing = f.origin[keep], f.spacing[keep]
        _save(os.path.join(HERE, f"wavelet_refine_{wav}.npz"), {"vor": sub.blocks}, sub,
              {"wavelet": wav, "subsample": int(wav != "CDF40") * 8, "n_total": int(f.blocks.shape[0])})
    key = {(int(t), int(l)): n for n, (t, l) in enumerate(zip(src.treecode, src.level))}
    for wav in ("CDF20", "CDF22", "CDF40", "CDF42", "CDF44", "CDF60", "CDF62"):
        f = read_wabbit_file(os.path.join(W, f"adaptive_{wav}", "vor_00200.h5"))
        idx = [key[(int(t), int(l))] for t, l in zip(f.treecode, f.level)]
        if wav in ("CDF20", "CDF40", "CDF60"):
            # interiors identical to the input (the stored extra point is a
            # synced ghost value and may differ)
            assert np.array_equal(f.blocks[:, :-1, :-1], src.blocks[idx][:, :-1, :-1])
            continue
        _save(os.path.join(HERE, f"wavelet_coarsen_{wav}.npz"), {"vor": f.blocks}, f, {"wavelet": wav})


def acm_cylinder():
    """TESTING/acm/acm_CDF44: adaptive penalised cylinder with a p-norm sponge
    (acm_cyl.ini: Bs=26, Jmax=6, CDF44, eps=1e-3 normalised, mask
    thresholding, max-level dealiasing).  The saved state at t = 0, 0.05 and
    0.1 (ux, uy, p and the mask chi), one fixture per time."""
    d = "/root/reference/TESTING/acm/acm_CDF44"
    params = dict(Bs=26, Jmax=6, Jmin=1, domain=20.0, wavelet="CDF44", eps=1e-3, eps_normalized=True,
                  eps_norm="Linfty", threshold_mask=True, force_maxlevel_dealiasing=True,
                  c_0=12.5, nu=0.0, gamma_p=0.0, u_mean_set=[0.0, -1.0, 0.0], C_eta=1.34e-3,
                  C_sponge=8.0e-3, L_sponge=2.0, p_sponge=8.0, sponge_type="p-norm",
                  x_cntr=[10.0, 10.0], R_cyl=0.5, C_smooth=1.5, CFL=1.5, CFL_eta=0.99,
                  time_max=0.1, write_time=0.05, tsave_stats=0.2, skew_symmetry=False,
                  scheme="FD_4th_central")
    for tag, name in (("000000000000", "t0"), ("000000050000", "t005"), ("000000100000", "t010")):
        fields, meta = _fields(d, ["ux", "uy", "p", "mask"], tag)
        _save(os.path.join(HERE, f"acm_cyl_{name}.npz"), fields, meta, params)


if __name__ == "__main__":
    vortex2d()
    taylor_green()
    wavelets()
    acm_cylinder()
