#!/usr/bin/env python3
"""Benchmark of the block time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c1|c3u]

A "step" is one RK4 time step (4 RK stages over every block of the forest):
device-side dt reduction and clamps, ghost sync and the fused RHS+RK stage
kernels, replayed from a hipGraph.  Inputs are resident in HBM before the
timed region starts.  The metric is block-RK-stages/s = 4 * N_blocks * K / t.

Workloads
  c1   2D ACM, periodic unit square, uniform level 4 (256 blocks), Bs=16,
       CDF40 ghosts (g=3, g_rhs=2), FD_4th_central, RK4, CFL 0.5, c0=20,
       nu=1e-3; seeded Fourier velocity (|k|<=8, amplitude |k|^-2, u_rms=1)
       and a Gaussian pressure pulse (sigma 0.05, amplitude 0.1).
  c3u  3D ACM, periodic unit cube, uniform level 4 (4096 blocks), Bs=24,
       CDF44 ghosts (g=6, g_rhs=2), FD_4th_central, c0=15, nu=1e-3; seeded
       solenoidal velocity (1<=|k|<=6, E(k)~k^-5/3, u_rms=1), p=0.  The
       block size, ghosts and physics of the 3D adaptive workload on a
       uniform grid sized to exceed the 256 MB last-level cache.

Printed: one JSON line with the contract fields plus "roofline" (dominant
kernel: the fused RHS+RK stage kernel, HIP-event timed on the library's
stream) and "cpu_baseline" (the oracle's C restatement on the host cores, a
bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def b_min(q, bs, dim, g_r, m):
    """Algorithmic bytes per block-RK-stage of a fused RK4 (36qN + 8qH + 8mN)."""
    N = bs ** dim
    H = (bs + 2 * g_r) ** dim - N
    return 36 * q * N + 8 * q * H + 8 * m * N


def workload(name):
    from wabbit_amd.params import ACMParams, TimeParams
    if name == "c1":
        return dict(name="c1-2d-periodic-J4-Bs16", dim=2, J=4, bs=16, g=3, g_rhs=2, L=1.0,
                    acm=ACMParams(dim=2, scheme=4, skew_symmetry=False, penalization=False,
                                  c_0=20.0, nu=1e-3, gamma_p=1.0, CFL=0.5,
                                  domain_size=[1.0, 1.0, 1.0]),
                    ic=dict(kmax=8, kmin=1, amp_exponent=-2.0), pressure=True, nsteps_cfg=100)
    if name == "c3u":
        return dict(name="c3u-3d-periodic-J4-Bs24", dim=3, J=4, bs=24, g=6, g_rhs=2, L=1.0,
                    acm=ACMParams(dim=3, scheme=4, skew_symmetry=False, penalization=False,
                                  c_0=15.0, nu=1e-3, gamma_p=1.0, CFL=0.5,
                                  domain_size=[1.0, 1.0, 1.0]),
                    ic=dict(kmax=6, kmin=1, energy_exponent=-5.0 / 3.0, solenoidal=True),
                    pressure=False, nsteps_cfg=200)
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(w, hvy, forest, seconds=10.0):
    """The oracle (C restatement, OpenMP) on the host cores for a bounded
    number of steps of the same workload."""
    from oracle import pyoracle as po
    from wabbit_amd.params import TimeParams, clamp_dt
    threads = min(os.cpu_count() or 1, 16)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    dim = w["dim"]
    run = po.EquidistantRun(po.make_params(w["acm"]), [w["bs"]] * dim, w["g"], w["g_rhs"], dim + 1,
                            w["J"], [w["L"]] * dim, hvy.copy(), forest.block_at())
    tp = TimeParams(time_max=1e9)
    t, n = 0.0, 0
    t0 = time.perf_counter()
    while True:
        dt = clamp_dt(run.dt(), t, tp)
        run.step(dt)
        t += dt
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 100000:
            break
    return dict(value=4 * forest.n * n / el, unit="block-RK-stages/s", cores=threads, kind="port",
                sample=f"{n} RK4 steps of the same workload ({forest.n} blocks), {el:.1f} s, "
                       f"OMP_NUM_THREADS={threads}")


def calibrate_copy(name, reps=10):
    """A plain copy of known size for scaling the FETCH_SIZE / WRITE_SIZE
    counters: wh_rk_combine with no terms copies the interiors of all blocks
    of hvy_block into work slot 1 (reads and writes q*N doubles per block)."""
    import ctypes

    from wabbit_amd import lib as L
    from wabbit_amd.forest import EquidistantForest
    w = workload(name)
    dim = w["dim"]
    forest = EquidistantForest(dim, w["J"], [w["bs"]] * dim, w["g"], [w["L"]] * dim)
    pool = L.Pool(dim, forest.bs, forest.g, dim + 1, forest.n, n_work_slots=2)
    pool.set_active(forest.active, forest.x0dx)
    pool.fill(L.WH_BLOCK, 1.0)
    for _ in range(reps):
        L.check(L.load().wh_rk_combine(pool.h, L.WH_WORK + 1, L.WH_BLOCK, 0, None, None))
    pool.synchronize()
    nbytes = forest.n * (dim + 1) * w["bs"] ** dim * 8
    print(json.dumps({"calibration": "interior copy", "launches": reps,
                      "read_bytes_per_launch": nbytes, "write_bytes_per_launch": nbytes}))
    pool.close()
    del ctypes


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="c1", choices=["c1", "c3u"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-paths", action="store_true", help="skip the per-path comparison")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain interior copy of known size (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.workload)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from wabbit_amd import lib as L
    from wabbit_amd.forest import EquidistantForest
    from wabbit_amd.inicond import acm_state
    from wabbit_amd.params import TimeParams
    from wabbit_amd.timestep import PATH_FUSED, PATH_PER_BLOCK, PATH_REFERENCE, AcmSolver

    w = workload(args.workload)
    steps = args.steps if args.steps is not None else w["nsteps_cfg"]
    dim = w["dim"]
    forest = EquidistantForest(dim, w["J"], [w["bs"]] * dim, w["g"], [w["L"]] * dim)
    print(f"# workload {w['name']}: {forest.n} blocks, building IC", file=sys.stderr, flush=True)
    hvy = acm_state(forest, seed=1234, pressure=w["pressure"], **w["ic"])
    tp = TimeParams(time_max=1e9)
    solver = AcmSolver(forest, w["acm"], tp, w["g_rhs"], path=PATH_FUSED)
    solver.load(hvy)
    pool = solver.pool

    try:
        import torch
        torch_sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is only used for the contract's extra barrier
        torch_sync = lambda: None  # noqa: E731

    # autotune the stage kernel for this (Bs, g, neq), then warm up (also
    # captures the graph)
    tuned, tuned_ms = pool.autotune(solver.acm_c, reps=3)
    solver.load(hvy)
    print(f"# autotune: variant {tuned} ({tuned_ms:.4f} ms per RHS launch)", file=sys.stderr,
          flush=True)
    solver.run_device(args.warmup, use_graph=True)
    pool.synchronize()
    torch_sync()
    t0 = time.perf_counter()
    solver.run_device(steps, use_graph=True)
    pool.synchronize()
    torch_sync()
    el = time.perf_counter() - t0
    value = 4 * forest.n * steps / el
    print(f"# graph path: {steps} steps in {el:.4f} s", file=sys.stderr, flush=True)

    # kernel timing pass (non-graph fused path, events on the library stream)
    pool.timing(True)
    solver.run_device(max(3, min(10, steps)), use_graph=False)
    pool.synchronize()
    k_ms, k_n = pool.kernel_time(0)
    s_ms, s_n = pool.kernel_time(1)
    pool.timing(False)
    q = dim + 1
    bpb = b_min(q, w["bs"], dim, w["g_rhs"], 0)
    achieved = bpb * forest.n / (k_ms * 1e-3) / 1e9
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                    kernel="k_rhs%dd<FD4,fused>" % dim, avg_launch_ms=round(k_ms, 5),
                    launches=k_n, bytes_per_block_stage=bpb, blocks_per_launch=forest.n,
                    sync_kernel_avg_ms=round(s_ms, 5), tuned_variant=tuned)

    paths = {"fused_graph": round(value, 1)}
    if not args.no_paths:
        for name, path, n in (("reference_structured", PATH_REFERENCE, 5),
                              ("plain_per_block", PATH_PER_BLOCK, 2)):
            s2 = AcmSolver(forest, w["acm"], tp, w["g_rhs"], path=path)
            s2.load(hvy)
            s2.step()
            s2.pool.synchronize()
            t1 = time.perf_counter()
            for _ in range(n):
                s2.step()
            s2.pool.synchronize()
            paths[name] = round(4 * forest.n * n / (time.perf_counter() - t1), 1)
            s2.pool.close()
        s3 = AcmSolver(forest, w["acm"], tp, w["g_rhs"], path=PATH_FUSED)
        s3.load(hvy)
        s3.run_device(2, use_graph=False)
        s3.pool.synchronize()
        t1 = time.perf_counter()
        s3.run_device(10, use_graph=False)
        s3.pool.synchronize()
        paths["fused_no_graph"] = round(4 * forest.n * 10 / (time.perf_counter() - t1), 1)
        s3.pool.close()

    cpu = None
    if not args.no_cpu:
        print("# cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(w, hvy, forest, seconds=10.0)

    out = {
        "metric": "block-RK-stages/s",
        "value": round(value, 1),
        "unit": "block-RK-stages/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": w["name"], "blocks": forest.n, "Bs": w["bs"], "dim": dim,
                   "g": w["g"], "g_rhs": w["g_rhs"], "scheme": "FD_4th_central", "rk": "RK4",
                   "neq": q},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "paths_block_RK_stages_per_s": paths,
    }
    print(json.dumps(out), flush=True)
    pool.close()


if __name__ == "__main__":
    main()
