"""GPU: the plain 2D workload (config 1) as configured, against the oracle.

Periodic unit square, uniform level 4 (256 blocks of Bs=16), CDF40 ghosts
(g=3, RHS ghosts 2), FD_4th_central non-skew, RK4 at CFL 0.5, c0=20,
nu=1e-3, the seeded Fourier-mode velocity (|k| <= 8, amplitudes ~ |k|^-2,
u_rms = 1) and the Gaussian pressure pulse; 100 steps on the device path that
keeps dt, its clamps and the time on the GPU and replays each step from a
hipGraph (wh_run_steps).  The oracle runs the same 100 steps on the host.

Bars: time after 100 steps equal to the oracle's; fields within relative L2
1e-9 (the kernels follow the oracle's order of operations, so the test also
reports whether they are bit for bit equal)."""
import numpy as np
import pytest

from oracle import pyoracle as po
from wabbit_amd import workloads as wl
from wabbit_amd.forest import EquidistantForest
from wabbit_amd.inicond import acm_state
from wabbit_amd.params import TimeParams, clamp_dt
from wabbit_amd.timestep import PATH_FUSED, AcmSolver

pytestmark = pytest.mark.gpu


def test_config1_100_steps_graph_path():
    w = wl.get("c1")
    dim, bs, g = w["dim"], w["bs"], w["g"]
    forest = EquidistantForest(dim, w["J"], [bs] * dim, g, [w["L"]] * dim)
    assert forest.n == 256
    hvy = acm_state(forest, seed=1234, pressure=w["pressure"], **w["ic"])
    tp = TimeParams(time_max=1e9)
    nsteps = w["nsteps_cfg"]
    assert nsteps == 100
    solver = AcmSolver(forest, w["acm"], tp, w["g_rhs"], path=PATH_FUSED)
    try:
        solver.load(hvy)
        solver.run_device(nsteps, use_graph=True)
        got = solver.state()
        t_gpu, it_gpu = solver.time, solver.iteration
    finally:
        solver.pool.close()
    run = po.EquidistantRun(po.make_params(w["acm"]), [bs] * dim, g, w["g_rhs"], dim + 1, w["J"],
                            [w["L"]] * dim, hvy.copy(), forest.block_at())
    t = 0.0
    for _ in range(nsteps):
        dt = clamp_dt(run.dt(), t, tp)
        run.step(dt)
        t += dt
    want = run.hvy
    inner = (slice(None), slice(None), 0, slice(g, g + bs), slice(g, g + bs))
    rel = float(np.linalg.norm(got[inner] - want[inner]) / np.linalg.norm(want[inner]))
    print(f"config 1 after {nsteps} steps: t={t_gpu!r} (oracle {t!r}), rel L2 {rel:.3e}, "
          f"bit for bit {np.array_equal(got[inner], want[inner])}", flush=True)
    assert it_gpu == nsteps
    assert t_gpu == t
    assert rel <= 1e-9
