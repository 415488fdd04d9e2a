"""GPU: one RK4 step driven from Fortran through the bind(C) module
(tests/fortran/rk4_step_driver.f90 over wabbit_amd/fortran/wabbit_hip.f90),
as WABBIT's timeStep_tree would call the library: a 16-block periodic 2D
forest (Bs=16, g=3, FD4 non-skew) in the host's hvy_block(nx,ny,nz,neq,Nb)
layout.  The block dt minimum equals the oracle's bit for bit, the stepped
field is within relative L2 1e-12 of the oracle's RK4 step and equal, bit
for bit, to the same step through the Python (ctypes) binding."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from test_fortran_binding import FLANG, build_driver
from wabbit_amd import lib as L
from wabbit_amd.forest import EquidistantForest
from wabbit_amd.inicond import acm_state
from wabbit_amd.params import ACMParams, TimeParams, butcher_rk4
from wabbit_amd.timestep import PATH_FUSED, AcmSolver

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not in this image")
def test_rk4_step_through_fortran_module(tmp_path):
    import subprocess
    bs, g = 16, 3
    forest = EquidistantForest(2, 2, [bs, bs], g, [1.0, 1.0])
    acm = ACMParams(dim=2, scheme=4, skew_symmetry=False, penalization=False, c_0=20.0, nu=1e-3,
                    gamma_p=1.0, CFL=0.5, domain_size=[1.0, 1.0, 1.0])
    hvy = acm_state(forest, seed=3)
    pt = forest.same_level_patches(2)
    (tmp_path / "params.bin").write_bytes(bytes(L.acm_params_c(acm)))
    (tmp_path / "butcher.bin").write_bytes(np.ascontiguousarray(butcher_rk4(), np.float64).tobytes())
    # C order (nb, neq, nz, ny, nx) is the Fortran order hvy_block(nx, ny, nz, neq, nb)
    (tmp_path / "hvy.bin").write_bytes(np.ascontiguousarray(hvy, np.float64).tobytes())
    (tmp_path / "x0dx.bin").write_bytes(np.ascontiguousarray(forest.x0dx, np.float64).tobytes())
    (tmp_path / "patches.bin").write_bytes(np.ascontiguousarray(pt, L.PATCH_DTYPE).tobytes())
    exe = build_driver(tmp_path, "rk4_step_driver.f90")
    print(f"running {exe.name} on {forest.n} blocks, {len(pt)} patches", flush=True)
    res = subprocess.run([str(exe), str(tmp_path), str(forest.n), str(len(pt)), str(bs), str(g)],
                         capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0
    dt_f = float(np.frombuffer((tmp_path / "dt.bin").read_bytes(), np.float64)[0])
    got = np.frombuffer((tmp_path / "out.bin").read_bytes(), np.float64).reshape(hvy.shape)

    run = po.EquidistantRun(po.make_params(acm), [bs, bs], g, 2, 3, 2, [1.0, 1.0], hvy.copy(), forest.block_at())
    dt_o = run.dt()
    run.step(dt_o)
    inner = (slice(None), slice(None), 0, slice(g, g + bs), slice(g, g + bs))
    want = run.hvy[inner]
    err = float(np.linalg.norm(got[inner] - want) / np.linalg.norm(want))
    print(f"dt fortran {dt_f!r} oracle {dt_o!r}; rel L2 vs oracle {err:.3e}", flush=True)
    assert dt_f == dt_o
    assert err <= 1e-12

    s = AcmSolver(forest, acm, TimeParams(time_max=10.0), g_rhs=2, path=PATH_FUSED)
    try:
        s.load(hvy)
        s.pool.rk_step(s.acm_c, s.butcher, 0.0, dt_f, PATH_FUSED)
        assert np.array_equal(s.state()[inner], got[inner])
    finally:
        s.pool.close()
