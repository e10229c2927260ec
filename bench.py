"""Benchmark: FMM particle-steps/s on a 1,048,576-particle Plummer sphere (one MI355X).

Workload (one "step" = one KDK leapfrog step = tree rebuild + moments + walk +
M2L/L2L + P2P/L2P + integration): N = 1,048,576 Plummer (seed 1, a = 1, M = 1,
G = 1, truncated at r = 20a), eps = 1e-3, theta = 0.5, leaf capacity 32,
Cartesian order p = 4 (M2L keeping every term |k| + |n| <= 5), FP32 force
kernels, FP64 positions and velocities, dt = 1e-3.  One force evaluation is
checked against the FP64 direct sum on 4096 random targets; then W warm-up
steps and K timed steps.  The plain FP32 port (same operator, untuned kernels)
is timed beside it for 3 steps.

Prints one JSON line.  The timed steps run without the library's per-phase
HIP events; `phases_ms` comes from a phase pass of --phase-steps steps after
them with the events on.  The library runs four streams: the moments and the
same-level M2L run beside the walk, P2P beside M2L and L2L, so `phases_ms`
overlap (`overlapped` names what each ran beside) and only `total` is wall
time.  The roofline entry is for the dominant kernel, the same-level M2L
(most of the step's FLOP), with the P2P kernel as `second`: both timed with
the library's HIP events on the stream each runs on, in a roofline pass of
--roofline-steps steps after the timed ones with every kernel of an
evaluation on one stream (kernels that share the GPU take longer than
alone; `in_timed_steps` gives their overlapped figures).  `traffic` comes
from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary.
"""
import argparse
import itertools
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 (vector and matrix), AMD published
HBM_PEAK_GBS = 8000.0
P2P_FLOP = 20              # per directed particle-particle interaction


def m2l_flop(p, q, terms="all", pot=True):
    """2 x #{kept (k, n): |k|, |n| <= p and |k| + |n| <= q ("all"), or
    |k| + |n| <= p, or |k| + |n| <= q with min(|k|, |n|) <= q - p ("low")}
    per directed cell pair (420 for p = q = 4; at p = 4, q = 5: 840 with
    every term, 600 with the low-degree set).  pot=False leaves out the
    output k = 0, which only the potential uses (a step's evaluations: 770
    and 530 at p = 4, q = 5)."""
    tri = lambda d: (d + 1) * (d + 2) // 2
    r = q if terms == "all" else q - p
    keep = lambda a, b: a + b <= p or (a + b <= q and min(a, b) <= r)
    return 2 * sum(tri(a) * tri(b) for a in range(0 if pot else 1, p + 1) for b in range(p + 1) if keep(a, b))


def ncoef(p):
    return (p + 1) * (p + 2) * (p + 3) // 6


def same_level_kernel(p, qc, num, den, lds_per_cu=160 * 1024):
    """Name of the library's same-level M2L kernel of a step (accelerations
    only) on one system at p <= 4: the pipelined form with 256 threads where
    two workgroups' tables (the same-level offsets that occur, a zero row and
    the offset map) fit one CU's LDS, 512 where one does."""
    q = qc & 15
    K = 2 * int(math.floor(math.sqrt(3.0) * den / num)) + 1
    acc = lambda x, y, z: num * num * (x * x + y * y + z * z) > 3 * den * den
    rows = set()
    for o in itertools.product(range(-K, K + 1), repeat=3):
        if acc(*o) and any(not acc(*[(c + v) // 2 for c, v in zip(cc, o)])
                           for cc in itertools.product((0, 1), repeat=3)):
            rows.add(tuple(abs(v) for v in o))
    lds = (len(rows) + 1) * (((ncoef(q) + 3) // 4) | 1) * 16 + ((K + 1) ** 3 * 2 + 15) // 16 * 16
    if lds > lds_per_cu:
        return f"m2l_lds_kernel<{p}, {qc}, false, 256, true, false>"
    return f"m2l_lds_kernel<{p}, {qc}, true, {256 if 2 * lds <= lds_per_cu else 512}, true, false>"


def measured_traffic(kernel, workload):
    """HBM bytes per launch from the committed FETCH_SIZE/WRITE_SIZE summary
    (rocprofv3 --pmc runs scaled by a copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_plummer1M_p4.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        fields = dict(kv.split("=", 1) for kv in line.split() if "=" in kv)
        if fields.get("kernel") == kernel.replace(" ", "") and fields.get("workload") == workload:
            return {"fetch_bytes": float(fields["fetch_bytes"]), "write_bytes": float(fields["write_bytes"]),
                    "source": "profiles/hbm_traffic_plummer1M_p4.md"}
    return None


def host_cpu():
    """Host CPU model (/proc/cpuinfo) and the number of cores this process
    may run on (what `nproc` prints)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count()
    return model, nproc


def cpu_baseline(sample_n, budget_s):
    """Reference NaiveSimulation (oracle/_ref, compiled from the reference's
    sources with -O2 -ffp-contract=off) timed on this host, 1 core; falls back
    to the oracle's FP32 restatement of the same loop when the reference
    build is absent.  Also times one step of the as-shipped -O0 build at
    N=4096 when it is there."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    def sample(n, seed=123):
        rng = np.random.default_rng(seed)
        P = np.zeros((n, 12), dtype=np.float32)
        P[:, 0:3] = rng.random((n, 3))
        P[:, 8] = 1.0 / n
        P[:, 9] = 1.0 / n
        return P

    n = sample_n
    P = sample(n)
    kind = "reference" if po.ref_lib() is not None else "port"
    run = po.naive_reference if kind == "reference" else po.naive_compat
    t0 = time.perf_counter()
    steps = 0
    while True:
        run(P, 1.0, 1e-3, 1)
        steps += 1
        if time.perf_counter() - t0 > budget_s * 0.25 or steps >= 3:
            break
    dt = (time.perf_counter() - t0) / steps
    pairs_per_s = n * (n - 1) / 2 / dt
    target_n = 1 << 20
    value = 2.0 * pairs_per_s / (target_n - 1)
    model, nproc = host_cpu()
    out = {"value": value, "unit": "particle-steps/s", "cores": 1, "kind": kind,
           "sample": f"NaiveSimulation::step at N={n} uniform, {steps} step(s) of {dt:.2f} s "
                     f"({pairs_per_s:.3g} pairs/s, -O2 -ffp-contract=off), extrapolated to N=1,048,576 "
                     f"as 2*pairs_per_s/(N-1)",
           "host_cpu": model, "nproc": nproc}
    # the build's own OpenMP FP64 direct sum (include/nbody/cpu_simulation.h) on every host core, one
    # force evaluation on the same sample: labelled "build oracle", not "reference"
    from nbody_amd.cpu import CpuSimulation
    xs = P[:, 0:3].astype(np.float64)
    ms_ = np.full(n, 1.0 / n)
    cpu = CpuSimulation.from_arrays(xs, np.zeros_like(xs), ms_, G=1.0, eps=1e-3)
    t0 = time.perf_counter()
    cpu.forces()
    d1 = time.perf_counter() - t0
    cpu.close()
    p1 = n * (n - 1) / d1  # directed pairs per second
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or nproc
    out["build_oracle"] = {"value": p1 / (target_n - 1), "unit": "particle-steps/s", "cores": threads,
                           "sample": f"the build's OpenMP FP64 direct sum, one force evaluation at N={n} in "
                                     f"{d1:.2f} s ({p1:.3g} directed pairs/s on {threads} threads), extrapolated "
                                     f"to N=1,048,576 as pairs_per_s/(N-1)"}
    if kind == "reference" and po.ref_lib(as_shipped=True) is not None:
        n0 = 4096
        t0 = time.perf_counter()
        po.naive_reference(sample(n0), 1.0, 1e-3, 1, as_shipped=True)
        d0 = time.perf_counter() - t0
        p0 = n0 * (n0 - 1) / 2 / d0
        out["as_shipped_O0"] = {"value": 2.0 * p0 / (target_n - 1), "unit": "particle-steps/s",
                                "sample": f"NaiveSimulation::step at N={n0} uniform, 1 step of {d0:.2f} s "
                                          f"({p0:.3g} pairs/s, built -O0 as shipped), extrapolated the same way"}
    return out


def copy_calibration(nbytes):
    """One device-to-device hipMemcpy of nbytes (the runtime's blit kernel):
    nbytes read and nbytes written, to scale the FETCH_SIZE/WRITE_SIZE
    counters of a profiled run."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for buf in (a, b):
        if hip.hipMalloc(ctypes.byref(buf), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


SUMMED = ["ms_m2l_same", "m2l_same_pairs", "m2l_same_launches", "ms_p2p", "p2p_interactions"]


def time_steps(sim, warmup, steps, serial=False, totals=None, timing=True):
    """warmup untimed steps, then `steps` timed ones (each step ends with a
    stream synchronisation inside the library), with the library's per-phase
    HIP events on (timing=False: off, which also lets captured steps replay);
    serial: every kernel of an evaluation on one stream, so that each event
    pair brackets one kernel running alone.  totals (a dict) receives the
    sums over the timed steps of the SUMMED statistics."""
    for _ in range(warmup):
        sim.step(1)
    sim.set_timing(timing, serial=serial)
    agg = {k: 0.0 for k in ["ms_tree", "ms_moments", "ms_walk", "ms_m2l", "ms_l2l", "ms_p2p", "ms_total"]}
    inter = m2l_pairs = split_pairs = p2p_pairs = 0
    if totals is not None:
        totals.update({k: 0 for k in SUMMED})
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step(1)
        st = sim.stats()
        for k in agg:
            agg[k] += st[k]
        inter += st["p2p_interactions"]
        m2l_pairs += st["m2l_pairs"]
        split_pairs += st["split_pairs"]
        p2p_pairs += st["p2p_pairs"]
        if totals is not None:
            for k in SUMMED:
                totals[k] += st[k]
    wall = time.perf_counter() - t0
    sim.set_timing(False)
    return agg, wall, inter, m2l_pairs, split_pairs, p2p_pairs, sim.stats()


def force_error(sim, n, ntargets, seed):
    """RMS (sqrt(sum|da|^2 / sum|a|^2)) and max relative error of one force
    evaluation against the library's FP64 direct sum on random targets."""
    acc, _ = sim.forces()
    tg = np.sort(np.random.default_rng(seed).choice(n, size=min(ntargets, n), replace=False))
    ad, _ = sim.direct_fp64(tg)
    d = np.linalg.norm(acc[tg].astype(np.float64) - ad, axis=1)
    ref = np.linalg.norm(ad, axis=1)
    return {"rms": float(np.sqrt(np.mean(d ** 2) / np.mean(ref ** 2))), "max": float(np.max(d / ref)),
            "targets": int(len(tg))}


def workload_uniform4096(nb, args):
    """Uniform cube N=4096 (seed 0), eps=1e-2, theta=0.5, cap 16, p=4, 10 KDK
    steps at dt=1e-3: the tuned and the plain FP32 port, and the reference CPU
    path (NaiveSimulation, its own force law) on one host core."""
    n, K = 4096, 10
    x, v, m = nb.ics.uniform(n, 0)
    res = {}
    for label, m2l, p2p, graph in [("tuned", nb.M2L_GROUPED, nb.P2P_LANE, 0),
                                   ("tuned_graph", nb.M2L_GROUPED, nb.P2P_LANE, 1),
                                   ("plain_port", nb.M2L_PLAIN, nb.P2P_LEAF, 0)]:
        cfg = nb.default_config(G=1.0, eps=1e-2, theta=0.5, dt=1e-3, leaf_capacity=16, order=4, m2l=m2l, p2p=p2p,
                                graph=graph)
        sim = nb.HipSimulation.from_arrays(x, v, m, m, config=cfg)
        err = force_error(sim, n, n, 0)
        agg, wall, *_ = time_steps(sim, 2, K, timing=False)
        st = sim.stats()
        sim.close()
        res[label] = {"particle_steps_per_s": n * K / wall, "ms_per_step": wall / K * 1e3, "force_err": err}
        if graph:
            res[label]["graph"] = {k: st[k] for k in ("graph_captures", "graph_replays", "graph_fallbacks")}
    # the value: the tuned kernels with captured steps (the fastest at this size)
    out = {"metric": "particle-steps/s, N=4096 uniform, 10 KDK steps; force error vs FP64 direct",
           "value": res["tuned_graph"]["particle_steps_per_s"], "unit": "particle-steps/s", "n_gpus": 1, "steps": K,
           "warmup": 2, "ms_per_step": res["tuned_graph"]["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
           "config": {"workload": "uniform_N4096_seed0_eps1e-2_theta0.5_cap16_p4_kdk_dt1e-3"}, "results": res}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        P = np.zeros((n, 12), dtype=np.float32)
        P[:, 0:3] = x
        P[:, 8] = m
        P[:, 9] = m
        kind = "reference" if po.ref_lib() is not None else "port"
        run = po.naive_reference if kind == "reference" else po.naive_compat
        P1, _ = run(P, 1.0, 1e-3, 1)
        a_ref = (P1[:, 4:7].astype(np.float64) - 0.0) / 1e-3
        cfg = nb.default_config(G=1.0, eps=1e-2, theta=0.5, leaf_capacity=16)
        sim = nb.HipSimulation.from_arrays(x, v, m, m, config=cfg)
        ad, _ = sim.direct_fp64(np.arange(n))
        sim.close()
        # the reference path's force law differs (separation (dx,dx,dx), no
        # softening): its error against the softened FP64 sum is O(1)
        d = np.linalg.norm(a_ref - ad, axis=1)
        t0 = time.perf_counter()
        run(P, 1.0, 1e-3, K)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n * K / dt, "unit": "particle-steps/s", "cores": 1, "kind": kind,
                               "sample": f"NaiveSimulation::step, N=4096 uniform, {K} steps in {dt:.2f} s",
                               "force_err": {"rms": float(np.sqrt(np.mean(d ** 2) / np.mean(np.sum(ad ** 2, 1)))),
                                             "max": float(np.max(d / np.linalg.norm(ad, axis=1)))}}
    return out


def workload_clustered(nb, args):
    """Two Plummer spheres of 4,194,304 particles (seeds 2, 3) at (+-5, 0, 0)
    approaching at 0.3 each; eps=5e-4, theta=0.4, p=6 (M2L |k|+|n| <= 6),
    leaf capacity 32, depth <= 21; per M2L precision (FP32 on the VALU, FP32
    and bf16x3 products on MFMA for the same-level pairs), 1 force evaluation
    (error on 8192 targets) plus 5 KDK steps at dt=5e-4.  `value` is the
    precision chosen by --m2l (default: FP32 VALU)."""
    n_each = args.n // 2 if args.n != 1 << 20 else 4194304
    x, v, m = nb.ics.clustered_pair(n_each)
    n = len(m)
    m = m / m.sum()
    K = 5
    precs = [("fp32_valu", nb.M2L_GROUPED), ("fp32_mfma", nb.M2L_MFMA_F32), ("bf16x3_mfma", nb.M2L_BF16X3)]
    sel = {"mfma_f32": "fp32_mfma", "bf16x3": "bf16x3_mfma"}.get(args.m2l, "fp32_valu")
    errs, ms_by = {}, {}
    for name, kind in precs:
        print(f"[clustered] M2L {name}", file=sys.stderr, flush=True)
        cfg = nb.default_config(G=1.0, eps=5e-4, theta=0.4, dt=5e-4, leaf_capacity=32, order=6, m2l_order=6,
                                max_depth=21, m2l=kind)
        sim = nb.HipSimulation.from_arrays(x, v, m, m, config=cfg)
        errs[name] = force_error(sim, n, 8192, 8192)
        res = time_steps(sim, 1, K)
        ms_by[name] = res[1] / K * 1e3
        if name == sel:
            agg, wall, inter, m2l_pairs, split_pairs, p2p_pairs, st = res
        sim.close()
    return {"metric": "particle-steps/s, N=8,388,608 clustered pair, theta=0.4, p=6; force error vs FP64 direct",
            "value": n * K / wall, "unit": "particle-steps/s", "n_gpus": 1, "steps": K, "warmup": 1,
            "ms_per_step": wall / K * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp32", "data": "synthetic",
            "config": {"workload": f"clustered_pair_2x{n_each}_seeds2_3_eps5e-4_theta0.4_cap32_p6_q6_kdk_dt5e-4"},
            "m2l_precision": sel, "force_err": errs, "ms_per_step_by_m2l": ms_by,
            "phases_ms": {k[3:]: agg[k] / K for k in agg},
            "tree": {"nodes": st["num_nodes"], "leaves": st["num_leaves"], "depth": st["depth"]},
            "m2l_pairs_per_step": m2l_pairs / K, "p2p_interactions_per_step": inter / K}


def workload_ensemble(nb, args):
    """256 independent Plummer systems (seeds 100-355), N log-uniform in
    [512, 65536], eps=1e-2, theta=0.5, p=4, one handle; the library's
    autotuner (nbh_autotune) picks the leaf capacity from 16, 24, 32, 48, 64,
    96, 128 and the P2P tile from 64 and 128 staged sources (2 force
    evaluations each); then 20 timed steps and the per-system force error on
    256 targets each."""
    x, v, m, off = nb.ics.ensemble(256, 100, 512, 65536)
    n = len(m)
    cfg = nb.default_config(G=1.0, eps=1e-2, theta=0.5, dt=1e-3, leaf_capacity=32, order=4)
    sim = nb.HipSimulation.from_arrays(x, v, m, m, config=cfg, sys_off=off)
    t0 = time.perf_counter()
    cap, chunk, tuned_ms, table = sim.autotune([16, 24, 32, 48, 64, 96, 128], [64, 128], evals=2)
    tune_s = time.perf_counter() - t0
    tune = {f"cap{c}_chunk{k}": ms for (c, k), ms in table.items()}
    acc, _ = sim.forces()
    rng = np.random.default_rng(256)
    errs = []
    tg_all = []
    for s in range(256):
        b, e = int(off[s]), int(off[s + 1])
        tg_all.append(np.sort(rng.choice(np.arange(b, e), min(256, e - b), replace=False)))
    tg = np.concatenate(tg_all)
    ad, _ = sim.direct_fp64(tg)
    pos = 0
    for t in tg_all:
        a, r = acc[t].astype(np.float64), ad[pos:pos + len(t)]
        pos += len(t)
        errs.append(float(np.sqrt(np.sum((a - r) ** 2) / np.sum(r ** 2))))
    K = 20
    _, wall, *_ = time_steps(sim, 2, K, timing=False)
    # phases from a pass with the library's per-phase events on
    PK = 5
    agg = time_steps(sim, 0, PK)[0]
    sim.close()
    # the same tuned configuration with captured steps (timing off)
    gcfg = nb.default_config(G=1.0, eps=1e-2, theta=0.5, dt=1e-3, leaf_capacity=cap, p2p_chunk=chunk, order=4,
                             graph=1)
    gsim = nb.HipSimulation.from_arrays(x, v, m, m, config=gcfg, sys_off=off)
    _, gwall, *_ = time_steps(gsim, 2, K, timing=False)
    gst = gsim.stats()
    gsim.close()
    return {"metric": "aggregate particle-steps/s, 256-system Plummer ensemble; per-system force error",
            "value": n * K / wall, "unit": "particle-steps/s", "n_gpus": 1, "steps": K, "warmup": 2,
            "ms_per_step": wall / K * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp32", "data": "synthetic",
            "config": {"workload": "ensemble_256_plummer_seeds100_355_N512_65536_eps1e-2_theta0.5_p4",
                       "n_total": n, "leaf_capacity": cap, "p2p_chunk": chunk},
            "autotune": {"ms_per_eval": tune, "chosen": {"leaf_capacity": cap, "p2p_chunk": chunk},
                         "seconds": tune_s},
            "graph": {"ms_per_step": gwall / K * 1e3, "particle_steps_per_s": n * K / gwall,
                      **{k: gst[k] for k in ("graph_captures", "graph_replays", "graph_fallbacks")}},
            "force_err": {"per_system_rms_median": float(np.median(errs)), "per_system_rms_max": float(np.max(errs))},
            "phases_ms": {k[3:]: agg[k] / PK for k in agg}, "phases_source": f"phase pass ({PK} steps)"}


def workload_adjoint(nb, args):
    """Plummer N=16384 (seed 4), eps=1e-2, theta=0.5, p=4, 200 KDK steps at
    dt=1e-3, checkpoints every 20: forward trajectory vs adjoint (reverse)
    time for the loss mean |x_N - x*|^2, x* evolved from velocities perturbed
    by sigma=1e-3 noise (seed 5); the whole reverse pass on the device
    (nbh_adjoint_forward / nbh_adjoint_backward)."""
    from nbody_amd.adjoint import DeviceTrajectory
    n, steps, dt = 16384, 200, 1e-3
    x0, v0, m = nb.ics.plummer(n, 4)
    cfg = nb.default_config(G=1.0, eps=1e-2, theta=0.5, dt=dt, leaf_capacity=16, order=4)
    rng = np.random.default_rng(5)
    adj = DeviceTrajectory(x0, v0 + 1e-3 * rng.standard_normal((n, 3)), m, steps, ckpt_every=20, config=cfg)
    xt, _ = adj.forward()           # the target snapshot (also warms the handle)
    adj.sim.set_state(x0, v0)
    xN, vN = adj.forward()
    t_fwd = adj.seconds["forward"]
    loss = float(np.sum((xN - xt) ** 2) / (3 * n))
    gx, gv, gm = adj.backward(2.0 * (xN - xt) / (3 * n), np.zeros_like(vN))
    t_bwd = adj.seconds["backward"]
    adj.close()
    return {"metric": "adjoint cost: reverse (gradient) time / forward trajectory time, N=16384, 200 steps",
            "value": t_bwd / t_fwd, "unit": "ratio", "n_gpus": 1, "steps": steps, "warmup": 1,
            "ms_per_step": t_fwd / steps * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp32", "data": "synthetic",
            "config": {"workload": "adjoint_plummer_N16384_seed4_eps1e-2_theta0.5_p4_cap16_200steps_ckpt20"},
            "forward_s": t_fwd, "backward_s": t_bwd, "loss": loss,
            "grad_norms": {"x0": float(np.linalg.norm(gx)), "v0": float(np.linalg.norm(gv)), "m": float(np.linalg.norm(gm))}}


def workload_cpu_fmm(nb, args):
    """The multi-core CPU comparator on the headline configuration: the FP64
    OpenMP FMM of CpuSimulation (method "fmm", include/nbody/cpu_fmm.h) on the
    same Plummer N=1,048,576 (seed 1), eps=1e-3, theta=0.5, p=4, cap 32, on
    every OpenMP thread of the host; one force evaluation (the force work of
    a KDK step) timed after one untimed, error on 4096 targets against the
    GPU's FP64 direct sum."""
    from nbody_amd.cpu import CpuSimulation
    n = args.n
    x, v, m = nb.ics.plummer(n, seed=1, a=1.0, M=1.0, G=1.0, rmax=20.0)
    cpu = CpuSimulation.from_arrays(x, v, m, m, time_step=1e-3, G=1.0, eps=1e-3, method="fmm", order=args.order,
                                    theta=0.5, leaf_capacity=32)
    cpu.forces()
    t0 = time.perf_counter()
    acc, _ = cpu.forces()
    wall = time.perf_counter() - t0
    phases = cpu.fmm_times()
    stats = cpu.fmm_stats()
    cpu.close()
    cfg = nb.default_config(G=1.0, eps=1e-3, theta=0.5, leaf_capacity=32, order=args.order)
    sim = nb.HipSimulation.from_arrays(x, v, m, m, config=cfg)
    tg = np.sort(np.random.default_rng(4096).choice(n, size=min(4096, n), replace=False))
    ad, _ = sim.direct_fp64(tg)
    sim.close()
    d = np.linalg.norm(acc[tg] - ad, axis=1)
    ref = np.linalg.norm(ad, axis=1)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    return {"metric": "particle-steps/s, N=1M Plummer, CPU FMM (FP64, OpenMP); force error vs FP64 direct",
            "value": n / wall, "unit": "particle-steps/s", "n_gpus": 0, "cores": threads, "steps": 1, "warmup": 1,
            "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp64", "data": "synthetic",
            "config": {"workload": f"cpu_fmm_plummer_N{n}_seed1_eps1e-3_theta0.5_cap32_p{args.order}"},
            "force_err": {"rms": float(np.sqrt(np.mean(d ** 2) / np.mean(ref ** 2))), "max": float(np.max(d / ref)),
                          "targets": int(len(tg))},
            "phases_s": phases, "tree": stats}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--m2l", default="grouped", choices=["plain", "grouped", "bf16x3", "mfma_f32"])
    ap.add_argument("--p2p", default="lane", choices=["leaf", "lane"])
    ap.add_argument("--order", type=int, default=4)
    ap.add_argument("--m2l-order", type=int, default=0,
                    help="M2L truncation q, order or order + 1 (0: order + 1): terms with |k|, |n| <= order and "
                         "|k| + |n| <= q, as --m2l-terms selects")
    ap.add_argument("--m2l-terms", default="all", choices=["all", "low"],
                    help="all: every term of degree <= q; low: beyond degree order only those with "
                         "min(|k|, |n|) <= q - order")
    ap.add_argument("--p2p-chunk", default="128", choices=["auto", "64", "128"],
                    help="P2P tile: source particles staged per wavefront round (auto: nbh_autotune picks 64 or "
                         "128 on the workload before the warm-up; on this workload it picks 128, by ~1%%)")
    ap.add_argument("--graph", action="store_true",
                    help="captured steps (nbh_config.graph = 1) in the timed steps (the phase pass after them "
                         "runs uncaptured steps, since per-phase events keep steps out of a graph)")
    ap.add_argument("--phase-timing", action="store_true",
                    help="the library's per-phase HIP events on in the timed steps (default: off, and phases_ms "
                         "comes from a phase pass of --phase-steps steps after them)")
    ap.add_argument("--no-phase-timing", action="store_true",
                    help="(the default; kept for old command lines)")
    ap.add_argument("--phase-steps", type=int, default=5,
                    help="steps of the phase pass (per-phase events on) after the timed steps")
    ap.add_argument("--serial", action="store_true",
                    help="run every kernel of an evaluation on one stream in the timed steps too")
    ap.add_argument("--roofline-steps", type=int, default=3,
                    help="steps of the serialized roofline pass after the timed steps (0: take the roofline "
                         "from the timed steps)")
    ap.add_argument("--cpu-sample", type=int, default=65536)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error-check", action="store_true")
    ap.add_argument("--workload", default="plummer1m",
                    choices=["plummer1m", "uniform4096", "clustered", "ensemble", "adjoint", "cpu_fmm"],
                    help="plummer1m: the headline; the others print the line of another configuration")
    ap.add_argument("--no-plain-compare", action="store_true",
                    help="skip timing the plain FP32 port (per-target M2L threads, per-leaf P2P) beside the tuned path")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="before the workload, copy a buffer of this many MiB device to device once: a kernel "
                         "of known traffic that scales the FETCH_SIZE/WRITE_SIZE counters of a profiled run")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import nbody_amd as nb

    if nb.device_count() < 1:
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib << 20)
    others = {"uniform4096": workload_uniform4096, "clustered": workload_clustered,
              "ensemble": workload_ensemble, "adjoint": workload_adjoint, "cpu_fmm": workload_cpu_fmm}
    if args.workload in others:
        print(json.dumps(others[args.workload](nb, args)), flush=True)
        return
    m2l = {"plain": nb.M2L_PLAIN, "grouped": nb.M2L_GROUPED, "bf16x3": nb.M2L_BF16X3, "mfma_f32": nb.M2L_MFMA_F32}[args.m2l]
    p2p = {"leaf": nb.P2P_LEAF, "lane": nb.P2P_LANE}[args.p2p]
    n = args.n
    p = args.order
    x, v, m = nb.ics.plummer(n, seed=1, a=1.0, M=1.0, G=1.0, rmax=20.0)
    cfg = nb.default_config(G=1.0, eps=1e-3, theta=0.5, dt=1e-3, leaf_capacity=32, order=p, m2l=m2l, p2p=p2p,
                            m2l_order=args.m2l_order,
                            m2l_terms={"all": nb.M2L_TERMS_ALL, "low": nb.M2L_TERMS_LOW}[args.m2l_terms],
                            p2p_chunk=0 if args.p2p_chunk == "auto" else int(args.p2p_chunk),
                            graph=1 if args.graph else 0)
    sim = nb.HipSimulation.from_arrays(x, v, m, m, config=cfg)
    tuned = None
    if args.p2p_chunk == "auto" and args.p2p == "lane":
        # the P2P tile shape (the leaf capacity is part of the workload)
        _, chunk, _, table = sim.autotune(None, [64, 128], evals=2)
        tuned = {"p2p_chunk": chunk, "ms_per_eval": {f"chunk{k}": ms for (_, k), ms in table.items()}}

    # one force evaluation, checked against the FP64 direct sum
    err = {}
    if not args.no_error_check:
        acc, _ = sim.forces()
        tg = np.random.default_rng(4096).choice(n, size=min(4096, n), replace=False)
        ad, _ = sim.direct_fp64(tg)
        d = np.linalg.norm(acc[tg].astype(np.float64) - ad, axis=1)
        ref = np.linalg.norm(ad, axis=1)
        err = {"force_rms_rel_err": float(np.sqrt(np.mean(d ** 2) / np.mean(ref ** 2))),
               "force_max_rel_err": float(np.max(d / ref)),
               "force_err_targets": int(len(tg))}

    K = args.steps
    tot = {}
    # the timed steps run without the library's per-phase HIP events (about
    # 1% of a step: each event is a marker in its stream); a phase pass after
    # them, streams as in the timed steps, gives phases_ms
    # (--serial needs them: the library serializes its streams in timing mode 2)
    timing = (args.phase_timing or args.serial) and not args.graph
    timed = time_steps(sim, args.warmup, K, serial=args.serial, totals=tot, timing=timing)
    gstats = {k: sim.stats()[k] for k in ("graph_captures", "graph_replays", "graph_fallbacks")}
    PK, ptot = K, tot
    pagg = timed[0]
    if not timing and args.phase_steps > 0:
        PK, ptot = args.phase_steps, {}
        pagg = time_steps(sim, 0, PK, serial=args.serial, totals=ptot, timing=True)[0]
    # the roofline pass: R more steps with every kernel of an evaluation on
    # one stream, so that each event pair times one kernel running alone (in
    # the timed steps the same-level M2L launches of successive levels, the
    # walk and P2P share the GPU and each takes longer than alone)
    R = args.roofline_steps
    rl = {}
    if R > 0:
        time_steps(sim, 0, R, serial=True, totals=rl)
    q = sim.m2l_order
    terms = sim.m2l_terms
    q_same = int(sim.stats()["m2l_same_order"]) or q
    sim.close()
    _, wall, inter, m2l_pairs, split_pairs, p2p_pairs, st = timed
    value = n * K / wall
    m2l_ms = pagg["ms_m2l"] / PK
    # the timed steps evaluate accelerations only (no potential terms)
    m2l_tf = m2l_flop(p, q, terms, pot=False) * m2l_pairs / K / (m2l_ms * 1e-3) / 1e12 if m2l_ms > 0 else 0.0
    workload = f"plummer_N{n}_seed1_eps1e-3_theta0.5_cap32_p{p}_kdk_dt1e-3"
    chunk = int(sim.config.p2p_chunk) or 64
    # (the variants of a step: accelerations only)
    p2p_kernel = f"p2p_staged_kernel<{p}, {chunk // 64}, false>" if args.p2p == "lane" else f"p2p_kernel<{p}, false>"
    qc = (q | 16) if (terms == "all" and p < q < 2 * p) else q
    mfma = args.m2l in ("bf16x3", "mfma_f32")
    same_kernel = (f"m2l_offset_kernel<{p}, {'true' if args.m2l == 'bf16x3' else 'false'}>" if mfma
                   else same_level_kernel(p, qc, 1, 2))
    # the same-level pairs take the dense operator (every |k|, |n| <= p) on
    # the matrix cores, the truncation q with the chosen term set on the VALU
    # (the matrix-core kernels apply the dense operator, potential row included)
    same_flop = m2l_flop(p, q_same, "all") if mfma else m2l_flop(p, q_same, terms, pot=False)

    def kernel_roofline(src, steps, kernel, flop_per_step, ms_per_step, launches_per_step, bound, tr):
        ach = flop_per_step / (ms_per_step * 1e-3) / 1e12 if ms_per_step > 0 else 0.0
        return {"bound": bound, "achieved": ach, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": ach / FP32_PEAK_TFLOPS,
                "traffic": (tr["fetch_bytes"] + tr["write_bytes"]) if tr else None,
                "traffic_source": tr["source"] if tr else None,
                "kernel": kernel, "timed_on": src, "steps": steps,
                "launches_per_step": launches_per_step, "ms_per_step": ms_per_step,
                "ms_per_launch": ms_per_step / launches_per_step if launches_per_step else None,
                "flop_per_launch": flop_per_step / launches_per_step if launches_per_step else None,
                "flop_per_step": flop_per_step}

    # the dominant kernel is the same-level M2L (most of the step's FLOP);
    # P2P is the second entry
    src, S, tt = ("roofline pass (serialized, after the timed steps)", R, rl) if R > 0 else \
                 (("timed steps" if PK == K and ptot is tot else "phase pass") +
                  (" (serialized)" if args.serial else " (streams overlapped)"), PK, ptot)
    same_launches = tt["m2l_same_launches"] / S
    roof = kernel_roofline(src, S, same_kernel, same_flop * tt["m2l_same_pairs"] / S, tt["ms_m2l_same"] / S,
                           same_launches, "mfma" if mfma else "fp32-valu",
                           measured_traffic(same_kernel, workload))
    roof.update({"phase": "same-level M2L (cells of one level, offsets tabled per level)",
                 "flop_per_pair": same_flop, "pairs_per_step": tt["m2l_same_pairs"] / S,
                 "m2l_same_order": q_same})
    roof["second"] = kernel_roofline(src, S, p2p_kernel, P2P_FLOP * tt["p2p_interactions"] / S,
                                     tt["ms_p2p"] / S, 1, "fp32-valu", measured_traffic(p2p_kernel, workload))
    roof["second"].update({"phase": "P2P near field", "flop_per_pair": P2P_FLOP,
                           "pairs_per_step": tt["p2p_interactions"] / S})
    # the same kernels in steps like the timed ones (streams overlapped unless
    # --serial), from the phase pass when the timed steps ran without events
    roof["in_timed_steps"] = {
        "source": "timed steps" if ptot is tot else f"phase pass ({PK} steps after the timed steps)",
        "same_level_m2l_ms_per_step": ptot["ms_m2l_same"] / PK,
        "same_level_m2l_achieved": same_flop * ptot["m2l_same_pairs"] / PK / (ptot["ms_m2l_same"] / PK * 1e-3)
        / 1e12 if ptot["ms_m2l_same"] > 0 else None,
        "p2p_ms_per_step": ptot["ms_p2p"] / PK,
        "p2p_achieved": P2P_FLOP * ptot["p2p_interactions"] / PK / (ptot["ms_p2p"] / PK * 1e-3) / 1e12
        if ptot["ms_p2p"] > 0 else None}
    # algorithmic bytes of the HBM-bound phases (each input read once, each
    # output written once), per step
    nodes, leaves = st["num_nodes"], st["num_leaves"]
    ncp = (ncoef(p) + 3) & ~3
    phase_bytes = {
        # bbox 24 B + keys 36 B + 8 radix passes x 24 B + tree passes 24 B +
        # state gather 2 x 68 B per particle; 36 B per node record
        "tree": 412.0 * n + 36.0 * nodes,
        # P2M 52 B per particle, leaf moments 4 ncp B, M2M reads children and
        # writes the parent (about 9 rows of 4 ncp B per internal node)
        "moments": 52.0 * n + 4.0 * ncp * leaves + 9 * 4.0 * ncp * (nodes - leaves),
        # 4 B per emitted M2L / P2P / split entry, 40 B of row records per node
        "walk": 4.0 * (m2l_pairs + p2p_pairs + split_pairs) / K + 40.0 * nodes,
    }
    phase_roofline = {}
    for ph, b in phase_bytes.items():
        ms = pagg["ms_" + ph] / PK
        gbs = b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        phase_roofline[ph] = {"bound": "hbm", "bytes": b, "ms": ms, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS}
    phase_roofline["m2l"] = {"bound": "fp32-valu", "flop_per_pair": m2l_flop(p, q, terms, pot=False), "ms": m2l_ms,
                             "achieved_TFLOPs": m2l_tf, "frac": m2l_tf / FP32_PEAK_TFLOPS,
                             "note": "all M2L pairs over the M2L phase's wall time (from the first same-level "
                                     "launch, beside the walk, to the last other-level kernel)"}
    out = {
        "metric": "particle-steps/s, N=1M Plummer, force rms err<=1e-4 vs FP64 direct; %roofline",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": workload,
                   "n": n, "order": p, "m2l_order": q, "m2l_terms": terms, "theta": 0.5, "leaf_capacity": 32,
                   "eps": 1e-3, "m2l": args.m2l, "p2p": args.p2p, "p2p_chunk": chunk, "state_dtype": "fp64",
                   "streams": "serialized" if args.serial else "overlapped",
                   "captured_steps": bool(args.graph)},
        "roofline": roof,
        "autotune": tuned,
        "graph": gstats if args.graph else None,
        "phases_ms": {k[3:]: pagg[k] / PK for k in pagg},
        "phases_source": "timed steps" if ptot is tot else f"phase pass ({PK} steps after the timed steps)",
        "overlapped": {} if args.serial else {"moments": "walk", "m2l": "walk+p2p", "p2p": "m2l+l2l"},
        "phase_roofline": phase_roofline,
        "m2l": {"pairs_per_step": m2l_pairs / K, "same_level_pairs_per_step": tot["m2l_same_pairs"] / K,
                "flop_per_pair": m2l_flop(p, q, terms, pot=False), "tflops": m2l_tf},
        "tree": {"nodes": nodes, "leaves": leaves, "depth": st["depth"]},
    }
    if not args.no_plain_compare and (args.m2l, args.p2p) != ("plain", "leaf"):
        # the plain all-FP32 port: one thread per target cell for M2L (tensor
        # built per pair), one wavefront per leaf for P2P, same operator
        pcfg = nb.default_config(G=1.0, eps=1e-3, theta=0.5, dt=1e-3, leaf_capacity=32, order=p,
                                 m2l=nb.M2L_PLAIN, p2p=nb.P2P_LEAF, m2l_order=args.m2l_order,
                                 m2l_terms=cfg.m2l_terms)
        psim = nb.HipSimulation.from_arrays(x, v, m, m, config=pcfg)
        _, pwall, _, _, _, _, _ = time_steps(psim, 1, 3)
        psim.close()
        out["plain_port"] = {"ms_per_step": pwall / 3 * 1e3, "speedup": (pwall / 3) / (wall / K)}
    out.update(err)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
