"""Shared helpers for the parity tests: fixture loading and grid assembly."""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_npz(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def fixture_to_hvy(fields, origin, spacing, Bs, g, dim):
    """Place file blocks (N, Bs+1, ...) into hvy layout (N, neq, nz, ny, nx)
    the way readHDF5vct_tree does (InputOutput.f90:675-678): the Bs+1 stored
    points land at ghost-inclusive indices g .. g+Bs."""
    n = fields[0].shape[0]
    nx = Bs + 2 * g
    nz = nx if dim == 3 else 1
    hvy = np.zeros((n, len(fields), nz, nx, nx), np.float64)
    for c, f in enumerate(fields):
        if dim == 2:
            hvy[:, c, 0, g:g + Bs + 1, g:g + Bs + 1] = f
        else:
            hvy[:, c, g:g + Bs + 1, g:g + Bs + 1, g:g + Bs + 1] = f
    return hvy


def block_index_map(origin, spacing, Bs, J, dim):
    """block_at[(k*nb + j)*nb + i] = file index of the block with grid index (i,j,k)."""
    nb = 2 ** J
    block_at = -np.ones(nb ** dim, np.int32)
    for n in range(origin.shape[0]):
        ijk = [int(round(origin[n, d] / (Bs * spacing[n, d]))) for d in range(dim)]
        if dim == 2:
            idx = ijk[1] * nb + ijk[0]
        else:
            idx = (ijk[2] * nb + ijk[1]) * nb + ijk[0]
        block_at[idx] = n
    assert (block_at >= 0).all()
    return block_at


def interior(hvy, Bs, g, dim):
    if dim == 2:
        return hvy[:, :, 0, g:g + Bs, g:g + Bs]
    return hvy[:, :, g:g + Bs, g:g + Bs, g:g + Bs]


def fixture_interior(fields, dim, Bs):
    if dim == 2:
        return np.stack([f[:, :Bs, :Bs] for f in fields], axis=1)
    return np.stack([f[:, :Bs, :Bs, :Bs] for f in fields], axis=1)


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def params_of(npz):
    return json.loads(str(npz["params_json"]))


# ---------------------------------------------------------------------------
# the reference's regression runs, as fixtures
# ---------------------------------------------------------------------------
GHOSTS = {2: (1, 1), 4: (3, 2), 6: (5, 3)}  # scheme -> (g of CDF{2,4,6}0, g_rhs)


def vortex2d_case(order):
    """3vorticesEquiFD{order}: returns dict with acm/time params, forest data,
    initial hvy (ghosts zero), golden interior output."""
    from wabbit_amd.params import ACMParams, TimeParams
    ic = load_npz("vortex2d_ic.npz")
    out = load_npz(f"vortex2d_fd{order}.npz")
    prm = params_of(out)
    Bs, J, dim = prm["Bs"], prm["Jmax"], 2
    g, g_rhs = GHOSTS[order]
    L = prm["L"]
    fields = [ic["field_ux"], ic["field_uy"], ic["field_p"]]
    acm = ACMParams(dim=2, scheme=order, skew_symmetry=True, penalization=False,
                    c_0=prm["c_0"], nu=prm["nu"], gamma_p=prm["gamma_p"], CFL=prm["CFL"],
                    domain_size=[L, L, 1.0])
    tp = TimeParams(time_max=prm["time_max"], write_method="fixed_time",
                    write_time=prm["write_time"])
    return dict(
        acm=acm, tp=tp, Bs=Bs, J=J, dim=dim, g=g, g_rhs=g_rhs, L=L,
        hvy=fixture_to_hvy(fields, ic["origin"], ic["spacing"], Bs, g, dim),
        origin=ic["origin"], spacing=ic["spacing"], treecode=ic["treecode"],
        time0=float(ic["time"]), it0=int(ic["iteration"]),
        time1=float(out["time"]), it1=int(out["iteration"]),
        golden=fixture_interior([out["field_ux"], out["field_uy"], out["field_p"]], dim, Bs),
    )


def tg3d_case(order):
    from wabbit_amd.params import ACMParams, TimeParams
    f = load_npz(f"tg3d_fd{order}.npz")
    prm = params_of(f)
    Bs, J, dim = prm["Bs"], prm["Jmax"], 3
    g, g_rhs = GHOSTS[order]
    L = prm["L"]
    names = ["ux", "uy", "uz", "p"]
    acm = ACMParams(dim=3, scheme=order, skew_symmetry=True, penalization=False,
                    c_0=prm["c_0"], nu=prm["nu"], gamma_p=prm["gamma_p"], CFL=prm["CFL"],
                    domain_size=[L, L, L])
    tp = TimeParams(time_max=prm["time_max"], write_method="fixed_time",
                    write_time=prm["write_time"], tsave_stats=prm["tsave_stats"])
    return dict(
        acm=acm, tp=tp, Bs=Bs, J=J, dim=dim, g=g, g_rhs=g_rhs, L=L,
        hvy=fixture_to_hvy([f[f"field_ic_{n}"] for n in names], f["origin"], f["spacing"], Bs, g, dim),
        origin=f["origin"], spacing=f["spacing"], treecode=f["treecode"],
        time0=prm["t0"], it0=prm["it0"], time1=prm["t1"], it1=prm["it1"],
        golden=fixture_interior([f[f"field_out_{n}"] for n in names], dim, Bs),
    )


def reorder_to_forest(case, forest):
    """Permute the file's block order into the forest's hvy order (by origin)."""
    bat = block_index_map(case["origin"], case["spacing"], case["Bs"], case["J"], case["dim"])
    # forest hvy id k has grid index unravel(k); file block bat[k]
    return case["hvy"][bat], bat


def oracle_run(case, nsteps=None):
    """Run the oracle from the case's initial state with the reference's
    main-loop rule; returns (hvy, time, iteration)."""
    from oracle import pyoracle as po
    from wabbit_amd.params import clamp_dt
    p = po.make_params(case["acm"])
    dim, Bs = case["dim"], case["Bs"]
    bat = block_index_map(case["origin"], case["spacing"], Bs, case["J"], dim)
    run = po.EquidistantRun(p, [Bs] * dim, case["g"], case["g_rhs"], dim + 1, case["J"],
                            [case["L"]] * dim, case["hvy"].copy(), bat)
    t, it = case["time0"], case["it0"]
    tp = case["tp"]
    n = 0
    while t < tp.time_max and it < tp.nt and (nsteps is None or n < nsteps):
        dt = clamp_dt(run.dt(), t, tp)
        run.step(dt)
        t += dt
        it += 1
        n += 1
    return run.hvy, t, it


def smooth_field(forest, neq, seed, amp=1.0):
    """Seeded smooth periodic field sampled at every point of every block
    (ghosts included), shape (n, neq, nz, ny, nx)."""
    rng = np.random.default_rng(seed)
    dim, g, bs = forest.dim, forest.g, forest.bs
    nx = bs[0] + 2 * g
    ny = bs[1] + 2 * g
    nz = bs[2] + 2 * g if dim == 3 else 1
    out = np.zeros((forest.n, neq, nz, ny, nx))
    L = forest.domain
    for k in range(forest.n):
        x0, dx = forest.x0dx[k, :3], forest.x0dx[k, 3:]
        xs = [x0[d] + (np.arange(n) - g) * dx[d] for d, n in zip(range(3), (nx, ny, nz))]
        if dim == 2:
            xs[2] = np.zeros(1)
        Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
        for c in range(neq):
            r2 = np.random.default_rng(seed * 1000 + c)
            v = np.zeros_like(X)
            for _ in range(4):
                kx, ky, kz = r2.integers(1, 4, size=3)
                ph = r2.uniform(0, 2 * np.pi)
                a = r2.uniform(0.2, 1.0)
                arg = 2 * np.pi * (kx * X / L[0] + ky * Y / L[1]) + ph
                if dim == 3:
                    arg = arg + 2 * np.pi * kz * Z / L[2]
                v += a * np.cos(arg)
            out[k, c] = amp * v
    del rng
    return out


def random_mask(forest, seed):
    """A 6-component mask: chi in [0,1], solid velocity, colour in {0,1,2},
    sponge in [0,1]."""
    rng = np.random.default_rng(seed)
    dim, g, bs = forest.dim, forest.g, forest.bs
    nx = bs[0] + 2 * g
    ny = bs[1] + 2 * g
    nz = bs[2] + 2 * g if dim == 3 else 1
    m = rng.uniform(0.0, 1.0, size=(forest.n, 6, nz, ny, nx))
    m[:, 4] = rng.integers(0, 3, size=m[:, 4].shape).astype(np.float64)
    return m


# ---------------------------------------------------------------------------
# random 2:1-balanced leaf forests and a host-side patch applier
# ---------------------------------------------------------------------------
def _tc(ijk, level, jmax, dim):
    from wabbit_amd.forest import treecode_from_ijk
    return treecode_from_ijk(list(ijk) + [0] * (3 - len(ijk)), level, jmax, dim)


def random_graded_forest(dim, jmin, jmax, seed, nref):
    """Leaves (level, ijk) of a periodic forest: uniform at jmin, nref random
    refinements, then refined until every pair of touching leaves (faces,
    edges, corners, across the periodic boundary) differs by at most one
    level.  Returns (treecodes, levels) sorted by (level, treecode)."""
    rng = np.random.default_rng(seed)
    nb = 2 ** jmin
    leaves = set()
    for idx in np.ndindex(*([nb] * dim)):
        leaves.add((jmin, tuple(int(v) for v in idx)))

    def children(l, ijk):
        out = []
        for c in np.ndindex(*([2] * dim)):
            out.append((l + 1, tuple(2 * ijk[d] + int(c[d]) for d in range(dim))))
        return out

    for _ in range(nref):
        cand = sorted(x for x in leaves if x[0] < jmax)
        if not cand:
            break
        l, ijk = cand[rng.integers(len(cand))]
        leaves.remove((l, ijk))
        leaves.update(children(l, ijk))
    ext = 2 ** jmax

    def box(l, ijk):
        s = 2 ** (jmax - l)
        return [(ijk[d] * s, (ijk[d] + 1) * s) for d in range(dim)]

    def touch(a, b):
        for d in range(dim):
            ok = False
            for sh in (-ext, 0, ext):
                if a[d][0] <= b[d][1] + sh and b[d][0] + sh <= a[d][1]:
                    ok = True
            if not ok:
                return False
        return True

    changed = True
    while changed:
        changed = False
        lst = sorted(leaves)
        boxes = {x: box(*x) for x in lst}
        for a in lst:
            for b in lst:
                if b[0] >= a[0] + 2 and touch(boxes[a], boxes[b]):
                    leaves.remove(a)
                    leaves.update(children(*a))
                    changed = True
                    break
            if changed:
                break
    lst = sorted(leaves)
    tcs = np.array([_tc(ijk, l, jmax, dim) for l, ijk in lst], np.int64)
    lev = np.array([l for l, _ in lst], np.int32)
    return tcs, lev


def apply_patches_host(hvy, patches, order):
    """Apply a ghost-patch table to hvy (n, neq, nz, ny, nx) in stage order;
    prediction from the oracle's whole-array prediction of the sender block
    taken after stages 1-2.  Test helper (host)."""
    from oracle import pyoracle as po
    out = hvy.copy()
    n, neq, nz, ny, nx = hvy.shape
    dim = 3 if nz > 1 else 2
    for stage in (1, 2, 3):
        snap = out.copy()
        pred_cache = {}
        for p in patches[patches["stage"] == stage]:
            r, s = int(p["recv"]), int(p["send"])
            rl, sl, nn = p["recv_lo"], p["send_lo"], p["n"]
            rsl = (slice(rl[2], rl[2] + nn[2]), slice(rl[1], rl[1] + nn[1]), slice(rl[0], rl[0] + nn[0]))
            if p["op"] == 0:
                ssl = (slice(sl[2], sl[2] + nn[2]), slice(sl[1], sl[1] + nn[1]), slice(sl[0], sl[0] + nn[0]))
                out[r][(slice(None),) + rsl] = snap[s][(slice(None),) + ssl]
            elif p["op"] == 1:
                ssl = tuple(slice(sl[d], sl[d] + 2 * (nn[d] - 1) + 1, 2) for d in (2, 1, 0))
                out[r][(slice(None),) + rsl] = snap[s][(slice(None),) + ssl]
            else:
                if s not in pred_cache:
                    fz = 2 * nz - 1 if dim == 3 else 1
                    fine = np.zeros((neq, fz, 2 * ny - 1, 2 * nx - 1))
                    for c in range(neq):
                        src = np.ascontiguousarray(snap[s, c])
                        dst = np.zeros((fz, 2 * ny - 1, 2 * nx - 1))
                        po.lib().or_prediction(po._d(src), nx, ny, nz, po._d(dst), order)
                        fine[c] = dst
                    pred_cache[s] = fine
                fsl = (slice(sl[2], sl[2] + nn[2]), slice(sl[1], sl[1] + nn[1]), slice(sl[0], sl[0] + nn[0]))
                out[r][(slice(None),) + rsl] = pred_cache[s][(slice(None),) + fsl]
    return out


def random_block_data(n, neq, dim, bs, g, seed):
    rng = np.random.default_rng(seed)
    nx = bs + 2 * g
    nz = nx if dim == 3 else 1
    return rng.standard_normal((n, neq, nz, nx, nx))


def smooth_blocks(tcs, lev, jmax, dim, bs, g, seed, neq):
    """Smooth periodic field with a sharp feature, sampled on every block
    (ghosts included) so the threshold keeps some blocks and drops others."""
    rng = np.random.default_rng(seed)
    n = tcs.shape[0]
    nz = bs + 2 * g if dim == 3 else 1
    out = np.zeros((n, neq, nz, bs + 2 * g, bs + 2 * g))
    cen = rng.uniform(0.2, 0.8, size=3)
    for k in range(n):
        l = int(lev[k])
        ijk = [0, 0, 0]
        for lv in range(1, l + 1):
            d = (int(tcs[k]) >> ((jmax - lv) * dim)) & ((1 << dim) - 1)
            ijk[0] = 2 * ijk[0] + ((d >> 1) & 1)
            ijk[1] = 2 * ijk[1] + (d & 1)
            ijk[2] = 2 * ijk[2] + ((d >> 2) & 1 if dim == 3 else 0)
        dx = 2.0 ** -l / bs
        xs = [(ijk[d] * bs + np.arange(bs + 2 * g) - g) * dx for d in range(3)]
        if dim == 2:
            xs[2] = np.zeros(1)
        Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
        r2 = (np.sin(np.pi * (X - cen[0])) ** 2 + np.sin(np.pi * (Y - cen[1])) ** 2 +
              (np.sin(np.pi * (Z - cen[2])) ** 2 if dim == 3 else 0.0))
        for c in range(neq):
            out[k, c] = np.exp(-r2 / (0.004 * (c + 1))) + 0.1 * np.cos(2 * np.pi * (X + c * Y))
    return out
