#!/usr/bin/env python3
"""Benchmark of the block time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c4|c2|c3|c5|c1|c3u]

A "step" is one time step of the workload's main loop.  The metric is
block-RK-stages/s: the RK4 stages times the blocks each stage swept, summed
over the K timed steps, divided by the wall time of those steps.  Inputs are
resident in HBM before the timed region starts.

c4 (default): the batched ensemble of 64 independent adaptive penalised
channel solves (2D ACM, CDF44, eps=1e-3, seeded cylinders, Jmax in
{4,5,6,7}, Bs in {16,24}), the largest configuration.  A step is one
main-loop iteration of every member: per member sync_ghosts_tree, refine
everywhere, the mask and dt; one RK4 step of all members (one fused stage
launch per RK stage over every block of every pool, replayed from a
hipGraph); per member adapt_tree.

c2: the adaptive penalised-cylinder channel alone (Bs=16, Jmin=2, Jmax=6).
c3: the adaptive 3D periodic workload.  c5: the serving stream.
c1 / c3u: uniform periodic grids (no adaptivity); a step is device-side
dt + four ghost syncs and fused RHS+RK stages, replayed from a hipGraph.

The workloads are defined in wabbit_amd/workloads.py.

Printed: one JSON line with the contract fields plus "roofline" (the fused
RHS+RK stage kernel timed back to back with HIP events on the library's
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
LLC_BYTES = 256 << 20  # MI355X last-level cache (Infinity Cache)


def frac_kind(working_set):
    """What the roofline fraction measures: an HBM fraction only where the
    stage's arrays exceed the last-level cache."""
    return "hbm" if working_set > LLC_BYTES else "effective bandwidth (working set within the 256 MB LLC)"


def stage_working_set(nblocks, bs, g, dim, q):
    """Bytes of the arrays a fused stage touches: input, accumulator, state
    and output, ghost-including blocks of q components."""
    return 4 * nblocks * q * (bs + 2 * g) ** dim * 8


def b_min(q, bs, dim, g_r, m):
    """Algorithmic bytes per block-RK-stage of a fused RK4 (36qN + 8qH + 8mN)."""
    N = bs ** dim
    H = (bs + 2 * g_r) ** dim - N
    return 36 * q * N + 8 * q * H + 8 * m * N


def workload(name):
    from wabbit_amd import workloads
    return workloads.get(name)


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


def _torch_sync():
    try:
        import torch
        if torch.cuda.is_available():
            return torch.cuda.synchronize
    except Exception:  # torch is only used for the contract's extra barrier
        pass
    return lambda: None


def cpu_baseline_adaptive(w, state, seconds=20.0):
    """The oracle's tree path (serial C) on the host for a bounded number of
    main-loop iterations of the adaptive workload from the same initial grid."""
    from oracle import pyoracle as po
    from wabbit_amd import workloads
    from wabbit_amd.params import butcher_rk4, clamp_dt
    tc, lv, hvy = state
    wo = po.wavelet(w["wavelet"], fd_max_size=2)
    if w["mask"] is not None:
        cfg = workloads.channel_adapt_config(w)
        tree = po.Tree(wo, w["dim"], w["jmin"], w["jmax"], [w["bs"]] * w["dim"], wo.g, tc, lv, hvy,
                       indicator=1, eps=w["eps"], groups=cfg.groups,
                       force_maxlevel_dealiasing=cfg.force_maxlevel_dealiasing,
                       eps_normalized=cfg.eps_normalized, threshold_mask=cfg.threshold_mask,
                       domain=w["domain"], mask=po.mask_params(**w["mask"]))
        mp = po.mask_params(**w["mask"])
    else:
        tree = po.Tree(wo, w["dim"], w["jmin"], w["jmax"], [w["bs"]] * w["dim"], wo.g, tc, lv, hvy,
                       indicator=1, eps=w["eps"], groups=[1] * (w["dim"] + 1), domain=w["domain"])
        mp = None
    ap = po.make_params(w["acm"])
    bt = butcher_rk4()
    t, n, stages = 0.0, 0, 0
    t0 = time.perf_counter()
    while True:
        tree.sync()
        tree.refine_everywhere()
        nb = tree.size()
        dt = clamp_dt(tree.dt(ap), t, w["tp"])
        tree.rk_step(ap, mp, bt, dt, 2)
        t += dt
        tree.adapt()
        n += 1
        stages += 4 * nb
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    tree.close()
    return dict(value=stages / el, unit="block-RK-stages/s", cores=1, kind="port",
                sample=f"{n} main-loop iterations of the same workload from its initial grid "
                       f"({len(tc)} blocks), {el:.1f} s, serial")


def run_adaptive(args, w):
    from wabbit_amd import lib as L
    from wabbit_amd import workloads
    from wabbit_amd.timestep import PATH_FUSED
    torch_sync = _torch_sync()
    steps = args.steps if args.steps is not None else w["nsteps_cfg"]
    warmup = args.warmup
    t0 = time.perf_counter()
    s = workloads.adaptive_solver(w, path=PATH_FUSED, nb_max=16384 if w["dim"] == 2 else 12288)
    print(f"# workload {w['name']}: initial grid {s.nblocks} blocks ({time.perf_counter() - t0:.1f} s)",
          file=sys.stderr, flush=True)
    init_state = s.state()
    for k in range(warmup):
        if k == warmup - 1:
            s.autotune_next_step()   # stage-kernel variant on a refined grid
        s.step()
    s.pool.synchronize()
    torch_sync()
    stages, nb_min, nb_max = 0, 1 << 30, 0
    t0 = time.perf_counter()
    for k in range(steps):
        s.step()
        stages += 4 * s.rk_blocks
        nb_min, nb_max = min(nb_min, s.rk_blocks), max(nb_max, s.rk_blocks)
        if (k + 1) % 100 == 0:
            print(f"# step {k + 1}: {s.rk_blocks} blocks, t={s.time:.5f}", file=sys.stderr, flush=True)
    s.pool.synchronize()
    torch_sync()
    el = time.perf_counter() - t0
    value = stages / el
    print(f"# adaptive loop: {steps} steps in {el:.3f} s", file=sys.stderr, flush=True)

    # kernel timing + per-phase pass over a few more steps
    nprof = max(3, min(20, steps))
    s.pool.timing(True)
    s.profile(True)
    st2 = 0
    for _ in range(nprof):
        s.step()
        st2 += 4 * s.rk_blocks
    k_ms, k_n = s.pool.kernel_time(0)
    phases = {k: round(v / nprof, 4) for k, v in s.phase_ms.items()}
    s.profile(False)
    s.pool.timing(False)
    # the stage kernel's launch duration: NREP RK4 stage sequences (4 NREP
    # launches) back to back on the last step's refined grid, one HIP event
    # pair around them (a per-launch event pair adds the dispatch gap to
    # short kernels)
    NREP = 5
    ks_ms = s.pool.time_stage_kernel(s.acm_c, NREP)
    q, dim = w["dim"] + 1, w["dim"]
    # m: mask components the RHS reads (2D penalised + sponge: chi, u_s, colour, sponge)
    m = 5 if w["mask"] is not None else 0
    bpb = b_min(q, w["bs"], dim, 2, m)
    bpl = s.rk_blocks
    achieved = bpb * bpl / (ks_ms * 1e-3) / 1e9
    ws = stage_working_set(bpl, w["bs"], s.g, dim, q)
    # traffic: the counter passes of the committed profile (they cannot be read inside the timed run)
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None, frac_of=frac_kind(ws),
                    working_set_bytes=ws,
                    kernel="k_rhs%dd<FD4,fused>" % dim, avg_launch_ms=round(ks_ms, 5), launches=4 * NREP,
                    bytes_per_block_stage=bpb, blocks_per_launch=bpl,
                    timing_pass_event_pair_ms=round(k_ms, 5), timing_pass_blocks_per_launch=round(st2 / max(k_n, 1), 1))

    if getattr(s, "tuned", None) is not None:
        roofline["tuned_variant"] = s.tuned[0]
        if w["dim"] == 3:
            roofline["tuned_zsplit"] = s.pool.get_tuning(L.WH_TUNE_K3D_ZSPLIT)
            roofline["tuned_waves"] = s.pool.get_tuning(L.WH_TUNE_K3D_WAVES)
            roofline["tuned_xcd"] = s.pool.get_tuning(L.WH_TUNE_K3D_XCD)
    cpu = None
    if not args.no_cpu:
        print("# cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline_adaptive(w, init_state, seconds=20.0)

    out = {
        "metric": "block-RK-stages/s",
        "value": round(value, 1),
        "unit": "block-RK-stages/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(el / steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": w["name"], "dim": dim, "Bs": w["bs"], "Jmin": w["jmin"], "Jmax": w["jmax"],
                   "wavelet": w["wavelet"], "eps": w["eps"], "scheme": "FD_4th_central", "rk": "RK4",
                   "neq": q, "initial_blocks": int(len(init_state[0])),
                   "rk_blocks_min": nb_min, "rk_blocks_max": nb_max, "time_end": round(s.time, 6)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "phase_ms_per_step": phases,
        # the RK step alone (the profiled pass's rk phase, synchronised)
        "stepper_only_block_RK_stages_per_s": round(st2 / nprof / (phases["rk"] * 1e-3), 1) if phases.get("rk") else None,
    }
    print(json.dumps(out), flush=True)
    s.close()


def cpu_baseline_ensemble(specs, states, seconds=15.0, threads=None):
    """The oracle's tree path (serial C per member) on the host cores: one
    host thread per member (the calls release the interpreter lock), each
    running its member's main loop from the member's initial grid until
    `seconds` have passed."""
    import threading

    from oracle import pyoracle as po
    from wabbit_amd.params import butcher_rk4, clamp_dt
    threads = threads or min(16, os.cpu_count() or 1, len(specs))
    stages = [0] * threads
    steps = [0] * threads
    t_end = [0.0]

    def work(i):
        spec, (tc, lv, hvy) = specs[i], states[i]
        cfg = spec.cfg
        wo = po.wavelet(cfg.wavelet, fd_max_size=2)
        mp = po.mask_params(**spec.mask)
        tree = po.Tree(wo, 2, cfg.jmin, cfg.jmax, [cfg.bs] * 2, wo.g, tc, lv, hvy, indicator=1, eps=cfg.eps,
                       groups=cfg.groups, eps_normalized=cfg.eps_normalized, threshold_mask=cfg.threshold_mask,
                       domain=spec.acm.domain_size, mask=mp)
        ap = po.make_params(spec.acm)
        t = 0.0
        while time.perf_counter() < t_end[0]:
            tree.sync()
            tree.refine_everywhere()
            nb = tree.size()
            dt = clamp_dt(tree.dt(ap), t, spec.tp)
            tree.rk_step(ap, mp, butcher_rk4(), dt, 2)
            t += dt
            tree.adapt()
            stages[i] += 4 * nb
            steps[i] += 1
        tree.close()

    t0 = time.perf_counter()
    t_end[0] = t0 + seconds
    th = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    el = time.perf_counter() - t0
    return dict(value=sum(stages) / el, unit="block-RK-stages/s", cores=threads, kind="port",
                sample=f"members 0-{threads - 1} of the ensemble from their initial grids, one host thread each "
                       f"running the oracle's serial C main loop for {el:.1f} s ({sum(steps)} member-steps)")


def run_ensemble(args):
    """c4: the 64-member ensemble of channel solves (seeded cylinders, Jmax
    in {4,5,6,7}, Bs in {16,24}).  Per step: per member sync_ghosts_tree,
    refine everywhere, mask (lanes of pools in parallel), then ONE RK4 step
    of all members (one fused stage launch per RK stage over every block of
    every pool, replayed from a hipGraph), then per member adapt_tree."""
    from wabbit_amd import workloads
    from wabbit_amd.batch import Batch
    from wabbit_amd.params import butcher_rk4
    torch_sync = _torch_sync()
    steps = args.steps if args.steps is not None else 200
    warmup = args.warmup
    specs = workloads.ensemble(args.members, nsteps=warmup + steps + 4)
    lanes = args.lanes
    tot, big = {}, {}
    for sp in specs:
        k = (2, sp.cfg.bs)
        tot[k] = tot.get(k, 0) + sp.slots
        big[k] = max(big.get(k, 0), sp.slots)
    # per lane: an even share plus room for the largest member (first-fit slack)
    slots = {k: lanes * (-(-tot[k] // lanes) + big[k]) for k in tot}
    t0 = time.perf_counter()
    b = Batch(slots, n_work_slots=3, lanes=lanes, parallel=not args.serial)
    solves = []
    for sp in sorted(specs, key=lambda x: -x.slots):      # largest first: packs the lanes
        s = b.admit(sp)
        if s is None:
            raise SystemExit(f"ensemble member {sp.tag} does not fit")
        solves.append(s)
    by_tag = {s.spec.tag: s for s in solves}
    ncpu = 0 if args.no_cpu else min(16, os.cpu_count() or 1)
    cpu_states = [by_tag[sp.tag].state() for sp in specs[:ncpu]]
    print(f"# ensemble: {len(specs)} members admitted in {lanes} lanes, "
          f"{sum(len(s.forest.blocks()[0]) for s in solves)} blocks ({time.perf_counter() - t0:.1f} s)",
          file=sys.stderr, flush=True)
    for _ in range(warmup):
        b.step()
    for g in b.groups.values():
        g.pool.synchronize()
    torch_sync()
    ph0 = dict(b.phase_s)
    t0 = time.perf_counter()
    stages = 0
    nb_min, nb_max = 1 << 40, 0
    for k in range(steps):
        stages += b.step()
        nb = sum(s.rk_blocks for s in solves)
        nb_min, nb_max = min(nb_min, nb), max(nb_max, nb)
        if (k + 1) % 10 == 0:
            print(f"# step {k + 1}: {nb} blocks", file=sys.stderr, flush=True)
    for g in b.groups.values():
        g.pool.synchronize()
    torch_sync()
    el = time.perf_counter() - t0
    value = stages / el
    phases = {k: round((b.phase_s.get(k, 0.0) - ph0.get(k, 0.0)) / steps * 1e3, 3) for k in b.phase_s}

    # the batched fused stage kernel alone: RK4 stage sequences back to back
    # on the last step's union grid (every block of every pool in one launch)
    NREP = 5
    ms, nblk = b.stage_kernel_roofline(NREP)[2]
    per_launch = sum(g.last_nact * b_min(3, g.key[1], 2, 2, 5 if g.key[3] else 0) for g in b.groups.values())
    achieved = per_launch / (ms * 1e-3) / 1e9
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                    kernel="k_rhs2d_multi<FD4,non-skew>", avg_launch_ms=round(ms, 5), launches=4 * NREP,
                    blocks_per_launch=nblk, bytes_per_launch=int(per_launch),
                    bytes_per_block_stage={f"Bs{bs}": b_min(3, bs, 2, 2, 5) for bs in (16, 24)},
                    working_set_bytes=int(per_launch), frac_of=frac_kind(per_launch))
    nblocks = sum(s.rk_blocks for s in solves)
    b.close()

    # the plain path (one launch per block and operation, as the reference
    # loops over blocks) against the batched path on the same 8 members
    plain = {}
    if not args.no_paths:
        sub = workloads.ensemble(8, nsteps=5)
        for name, path in (("batched", 2), ("per_block", 1)):
            bb = Batch({(2, 16): 40000, (2, 24): 40000}, n_work_slots=5)
            for sp in sub:
                bb.admit(sp)
            bb.step(path)
            for g in bb.groups.values():
                g.pool.synchronize()
            t1 = time.perf_counter()
            st = bb.step(path) + bb.step(path)
            for g in bb.groups.values():
                g.pool.synchronize()
            plain[name] = round(st / (time.perf_counter() - t1), 1)
            bb.close()
    cpu = None
    if ncpu:
        print("# cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline_ensemble(specs[:ncpu], cpu_states, seconds=15.0, threads=ncpu)
    out = {
        "metric": "block-RK-stages/s", "value": round(value, 1), "unit": "block-RK-stages/s", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": round(el / steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"c4-ensemble-{len(specs)}x-channel-J4-7-Bs16-24", "members": len(specs),
                   "lanes": lanes, "host_threads": "serial" if args.serial else "one per pool",
                   "blocks_per_step_min": nb_min, "blocks_per_step_max": nb_max,
                   "blocks_at_end": nblocks, "wavelet": "CDF44", "scheme": "FD_4th_central", "rk": "RK4"},
        "roofline": roofline, "cpu_baseline": cpu,
        "paths_block_RK_stages_per_s_8_members": plain,
        "phase_ms_per_step": phases,
    }
    print(json.dumps(out), flush=True)


def run_serving(args):
    """c5: a seeded Poisson stream of requests (70% 2D channel, 30% 3D
    periodic) on shared pools with continuous batching; the first
    --requests of the stream."""
    from wabbit_amd import workloads
    from wabbit_amd.serve import Request, Server
    from wabbit_amd.batch import budget_slots, pool_key, slot_bytes
    mix = workloads.serving_mix(n=args.requests, rate=args.rate)
    reqs = [Request(rid, spec, t) for rid, spec, t in mix]
    # every pool created at start-up within 60% of the device memory (the
    # reference's --memory): one pool per block shape and lane, so admission
    # only takes free slots; the 3D pool gets most of the budget (a 3D Bs=24
    # slot holds ~7.5 MB, a 2D one 0.13-0.22 MB)
    lanes = args.lanes
    keys = sorted({pool_key(r.spec) for r in reqs})
    from wabbit_amd import lib as L
    total = L.device_memory()[1]
    weights = {k: (0.78 if k[0] == 3 else 0.13 if k[1] == 24 else 0.09) for k in keys}
    slots = budget_slots(keys, weights, 0.6 * total, lanes)
    for k in keys:
        need = max(r.spec.slots for r in reqs if pool_key(r.spec) == k)
        if slots[(k[0], k[1])] // lanes < need:
            raise SystemExit(f"c5: a lane's share of the {k[0]}D Bs{k[1]} pool ({slots[(k[0], k[1])] // lanes} slots) "
                             f"is smaller than its largest request ({need}); use fewer lanes")
    pool_bytes = sum(slots[(k[0], k[1])] * slot_bytes(k) for k in keys)
    print(f"# c5: {len(keys)} pool shapes x {lanes} lanes, {pool_bytes / 1e9:.1f} GB of {total / 1e9:.1f} GB "
          f"({100 * pool_bytes / total:.0f} %) preallocated", file=sys.stderr, flush=True)
    srv = Server(slots, n_work_slots=3, lanes=lanes, prealloc=keys)
    st = srv.run(reqs, log=lambda m: print("# " + m, file=sys.stderr, flush=True))
    srv.close()
    out = {
        "metric": "solves/s", "value": round(st.solves_per_s, 4), "unit": "solves/s", "n_gpus": 1,
        "steps": None, "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"c5-serving-{len(reqs)}req-rate{args.rate}", "requests": len(reqs),
                   "lanes": args.lanes, "pool_slots": {f"{k[0]}d-Bs{k[1]}": v for k, v in slots.items()},
                   "preallocated_bytes": int(pool_bytes), "device_bytes": int(total)},
        "p50_latency_s": round(st.p50, 4), "p99_latency_s": round(st.p99, 4),
        "mean_queue_wait_s": round(st.mean_wait, 4), "completed": st.completed, "failed": st.failed,
        "wall_s": round(st.wall, 3), "block_RK_stages_per_s": round(st.block_rk_stages / st.wall, 1),
        "max_concurrent": st.max_concurrent,
        "failures": [dict(tag=r["tag"], error=r["error"]) for r in st.per_request if r["error"]],
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c4", choices=["c4", "c2", "c3", "c5", "c1", "c3u"])
    ap.add_argument("--members", type=int, default=64, help="c4: ensemble members")
    ap.add_argument("--requests", type=int, default=256, help="c5: requests taken from the stream")
    ap.add_argument("--rate", type=float, default=4.0, help="c5: Poisson arrival rate (1/s)")
    ap.add_argument("--lanes", type=int, default=None,
                    help="c4/c5: independent pool sets (every pool steps its solves on a host thread of its "
                         "own; default 2 -- c5's 3D pool share per lane must hold a Jmax-5 request within "
                         "60%% of HBM)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--serial", action="store_true",
                    help="c4: step the pools one after the other on the calling thread (the same device work; "
                         "for kernel tracers)")
    ap.add_argument("--no-paths", action="store_true", help="skip the per-path comparison")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain interior copy of known size (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.workload if args.workload in ("c1", "c3u") else "c1")
        return
    if args.gpus != 1:
        print(f"bench.py: the benchmark runs on one GPU; --gpus {args.gpus} is not supported "
              "(there is no multi-GPU decomposition)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "c4":
        args.lanes = args.lanes or 2
        run_ensemble(args)
        return
    if args.workload == "c5":
        args.lanes = args.lanes or 2
        run_serving(args)
        return
    w = workload(args.workload)
    if w["kind"] in ("adaptive", "adaptive3d"):
        run_adaptive(args, w)
        return

    from wabbit_amd import lib as L
    from wabbit_amd.forest import EquidistantForest
    from wabbit_amd.inicond import acm_state
    from wabbit_amd.params import TimeParams
    from wabbit_amd.timestep import PATH_FUSED, PATH_PER_BLOCK, PATH_REFERENCE, AcmSolver

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
    # the stage kernel back to back (NREP RK4 stage sequences, one HIP
    # event pair around their 4 NREP launches)
    NREP = 5
    ks_ms = pool.time_stage_kernel(solver.acm_c, NREP)
    q = dim + 1
    bpb = b_min(q, w["bs"], dim, w["g_rhs"], 0)
    achieved = bpb * forest.n / (ks_ms * 1e-3) / 1e9
    ws = stage_working_set(forest.n, w["bs"], w["g"], dim, q)
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None, frac_of=frac_kind(ws),
                    working_set_bytes=ws,
                    kernel="k_rhs%dd<FD4,fused>" % dim, avg_launch_ms=round(ks_ms, 5),
                    launches=4 * NREP, bytes_per_block_stage=bpb, blocks_per_launch=forest.n,
                    timing_pass_event_pair_ms=round(k_ms, 5),
                    sync_kernel_avg_ms=round(s_ms, 5) if s_n else None, tuned_variant=tuned)
    if dim == 2:
        # 1: the stages gather their halos from the neighbours' interiors
        # (no ghost-fill launches: sync_kernel_avg_ms is then null)
        roofline["halo_gather"] = pool.get_tuning(L.WH_TUNE_GATHER)
    if dim == 3:
        roofline["tuned_zsplit"] = pool.get_tuning(L.WH_TUNE_K3D_ZSPLIT)
        roofline["tuned_waves"] = pool.get_tuning(L.WH_TUNE_K3D_WAVES)
        roofline["tuned_xcd"] = pool.get_tuning(L.WH_TUNE_K3D_XCD)

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
