"""GPU: the fused RK stages reading each block's halo from its neighbours'
interiors (WH_TUNE_GATHER, a uniform 2D grid whose RHS ghost fill is all
same-level copies) against the same steps with the ghost fill launched
first: block interiors bit for bit equal after every step, for the device
time loop replayed from a hipGraph and for single steps."""
import numpy as np
import pytest

from wabbit_amd import lib as L
from wabbit_amd.forest import EquidistantForest
from wabbit_amd.params import ACMParams, TimeParams, butcher_rk4

pytestmark = pytest.mark.gpu

BS, G, J = 16, 3, 3


def _run(gather, use_graph, nsteps):
    f = EquidistantForest(2, J, [BS, BS], G, [1.0, 1.0])
    rng = np.random.default_rng(5)
    p = L.Pool(2, [BS, BS], G, 3, f.n, n_work_slots=3)
    try:
        p.set_tuning(L.WH_TUNE_GATHER, gather)
        assert p.get_tuning(L.WH_TUNE_GATHER) == gather
        p.upload(L.WH_BLOCK, 0, 0.2 * rng.standard_normal((f.n,) + p.block_shape()))
        p.set_active(f.active.astype(np.int32), f.x0dx)
        p.set_patches(0, f.same_level_patches(2))
        acm = ACMParams(dim=2, scheme=4, skew_symmetry=False, penalization=False, c_0=20.0, nu=1e-3,
                        gamma_p=1.0, CFL=0.5, domain_size=[1.0, 1.0, 1.0])
        ac = L.acm_params_c(acm)
        if use_graph:
            p.time_set(0.0, 0)
            p.run_steps(ac, butcher_rk4(), L.time_params_c(TimeParams(time_max=1e9)), nsteps, use_graph=True)
        else:
            for _ in range(nsteps):
                p.rk_step(ac, butcher_rk4(), 0.0, 1e-4, 2)
        out = p.download(L.WH_BLOCK, 0, f.n)
        return out[:, :, :, G:G + BS, G:G + BS]
    finally:
        p.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_halo_gather_matches_ghost_fill(use_graph):
    on, off = _run(1, use_graph, 6), _run(0, use_graph, 6)
    print(f"gather vs ghost fill, graph={use_graph}: max |diff| {np.abs(on - off).max()}", flush=True)
    assert np.array_equal(on, off)
