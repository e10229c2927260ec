"""Benchmark: FMM particle-steps/s on a 1,048,576-particle Plummer sphere (one MI355X).

Workload (one "step" = one KDK leapfrog step = tree rebuild + moments + walk +
M2L/L2L + P2P/L2P + integration): N = 1,048,576 Plummer (seed 1, a = 1, M = 1,
G = 1, truncated at r = 20a), eps = 1e-3, theta = 0.5, leaf capacity 32,
Cartesian order p = 4 (M2L keeping |k| + |n| <= 5), FP32 force kernels, FP64
positions and velocities, dt = 1e-3.  One force evaluation is checked against
the FP64 direct sum on 4096 random targets; then W warm-up steps and K timed
steps.  The plain FP32 port (same operator, untuned kernels) is timed beside it
for 3 steps.

Prints one JSON line.  The roofline entry is for the dominant kernel (P2P),
timed with HIP events on the library's stream inside the timed region; its
`traffic` comes from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 (vector and matrix), AMD published
HBM_PEAK_GBS = 8000.0
P2P_FLOP = 20              # per directed particle-particle interaction


def m2l_flop(p, q):
    """2 x #{(k, n): |k|, |n| <= p, |k| + |n| <= q} per directed cell pair
    (420 for p = q = 4, 840 for p = 4, q = 5)."""
    tri = lambda d: (d + 1) * (d + 2) // 2
    return 2 * sum(tri(a) * tri(b) for a in range(p + 1) for b in range(p + 1) if a + b <= q)


def ncoef(p):
    return (p + 1) * (p + 2) * (p + 3) // 6


def measured_traffic(kernel, workload):
    """HBM bytes per launch from the committed FETCH_SIZE/WRITE_SIZE summary
    (rocprofv3 --pmc runs scaled by a copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_plummer1M_p4.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        fields = dict(kv.split("=", 1) for kv in line.split() if "=" in kv)
        if fields.get("kernel") == kernel and fields.get("workload") == workload:
            return {"fetch_bytes": float(fields["fetch_bytes"]), "write_bytes": float(fields["write_bytes"]),
                    "source": "profiles/hbm_traffic_plummer1M_p4.md"}
    return None


def cpu_baseline(sample_n, budget_s):
    """Reference NaiveSimulation (oracle/_ref, compiled from the reference's
    sources) timed on this host, 1 core; falls back to the oracle's FP32
    restatement of the same loop when the reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    rng = np.random.default_rng(123)
    n = sample_n
    P = np.zeros((n, 12), dtype=np.float32)
    P[:, 0:3] = rng.random((n, 3))
    P[:, 8] = 1.0 / n
    P[:, 9] = 1.0 / n
    kind = "reference" if po.ref_lib() is not None else "port"
    run = po.naive_reference if kind == "reference" else po.naive_compat
    t0 = time.perf_counter()
    steps = 0
    while True:
        run(P, 1.0, 1e-3, 1)
        steps += 1
        if time.perf_counter() - t0 > budget_s * 0.5 or steps >= 3:
            break
    dt = (time.perf_counter() - t0) / steps
    pairs_per_s = n * (n - 1) / 2 / dt
    target_n = 1 << 20
    value = 2.0 * pairs_per_s / (target_n - 1)
    return {"value": value, "unit": "particle-steps/s", "cores": 1, "kind": kind,
            "sample": f"NaiveSimulation::step at N={n} uniform, {steps} step(s) of {dt:.2f} s "
                      f"({pairs_per_s:.3g} pairs/s, -O2), extrapolated to N=1,048,576 as 2*pairs_per_s/(N-1)"}


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


def time_steps(sim, warmup, steps):
    """warmup untimed steps, then `steps` timed ones (each step ends with a
    stream synchronisation inside the library)."""
    for _ in range(warmup):
        sim.step(1)
    sim.set_timing(True)
    agg = {k: 0.0 for k in ["ms_tree", "ms_moments", "ms_walk", "ms_m2l", "ms_l2l", "ms_p2p", "ms_total"]}
    inter = m2l_pairs = split_pairs = p2p_pairs = 0
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
    wall = time.perf_counter() - t0
    return agg, wall, inter, m2l_pairs, split_pairs, p2p_pairs, sim.stats()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--m2l", default="grouped", choices=["plain", "grouped", "bf16x3"])
    ap.add_argument("--p2p", default="lane", choices=["leaf", "lane"])
    ap.add_argument("--order", type=int, default=4)
    ap.add_argument("--cpu-sample", type=int, default=32768)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error-check", action="store_true")
    ap.add_argument("--no-plain-compare", action="store_true",
                    help="skip timing the plain FP32 port (per-target M2L threads, per-leaf P2P) beside the tuned path")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="before the workload, copy a buffer of this many MiB device to device once: a kernel "
                         "of known traffic that scales the FETCH_SIZE/WRITE_SIZE counters of a profiled run")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the sharding design is in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import nbody_amd as nb

    if nb.device_count() < 1:
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib << 20)
    m2l = {"plain": nb.M2L_PLAIN, "grouped": nb.M2L_GROUPED, "bf16x3": nb.M2L_BF16X3}[args.m2l]
    p2p = {"leaf": nb.P2P_LEAF, "lane": nb.P2P_LANE}[args.p2p]
    n = args.n
    p = args.order
    x, v, m = nb.ics.plummer(n, seed=1, a=1.0, M=1.0, G=1.0, rmax=20.0)
    cfg = nb.default_config(G=1.0, eps=1e-3, theta=0.5, dt=1e-3, leaf_capacity=32, order=p, m2l=m2l, p2p=p2p)
    sim = nb.HipSimulation.from_arrays(x, v, m, m, config=cfg)

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
    timed = time_steps(sim, args.warmup, K)
    q = sim.m2l_order
    sim.close()
    agg, wall, inter, m2l_pairs, split_pairs, p2p_pairs, st = timed
    value = n * K / wall
    p2p_ms = agg["ms_p2p"] / K
    p2p_flop = P2P_FLOP * inter / K
    achieved = p2p_flop / (p2p_ms * 1e-3) / 1e12
    m2l_ms = agg["ms_m2l"] / K
    m2l_tf = m2l_flop(p, q) * m2l_pairs / K / (m2l_ms * 1e-3) / 1e12 if m2l_ms > 0 else 0.0
    workload = f"plummer_N{n}_seed1_eps1e-3_theta0.5_cap32_p{p}_kdk_dt1e-3"
    kernel = ("p2p_staged_kernel" if args.p2p == "lane" else "p2p_kernel") + f"<{p}>"
    tr = measured_traffic(kernel, workload)
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
        ms = agg["ms_" + ph] / K
        gbs = b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        phase_roofline[ph] = {"bound": "hbm", "bytes": b, "ms": ms, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS}
    phase_roofline["m2l"] = {"bound": "fp32-valu", "flop_per_pair": m2l_flop(p, q), "ms": m2l_ms,
                             "achieved_TFLOPs": m2l_tf, "frac": m2l_tf / FP32_PEAK_TFLOPS}
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
                   "n": n, "order": p, "m2l_order": q, "theta": 0.5, "leaf_capacity": 32, "eps": 1e-3,
                   "m2l": args.m2l, "p2p": args.p2p, "state_dtype": "fp64"},
        "roofline": {"bound": "mfma", "engine": "fp32-valu (P2P kernel)", "achieved": achieved,
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": (tr["fetch_bytes"] + tr["write_bytes"]) if tr else None,
                     "traffic_source": tr["source"] if tr else None,
                     "kernel": kernel, "ms_per_launch": p2p_ms, "flop_per_launch": p2p_flop},
        "phases_ms": {k[3:]: agg[k] / K for k in agg},
        "phase_roofline": phase_roofline,
        "m2l": {"pairs_per_step": m2l_pairs / K, "tflops": m2l_tf},
        "tree": {"nodes": nodes, "leaves": leaves, "depth": st["depth"]},
    }
    if not args.no_plain_compare and (args.m2l, args.p2p) != ("plain", "leaf"):
        # the plain all-FP32 port: one thread per target cell for M2L (tensor
        # built per pair), one wavefront per leaf for P2P, same operator
        pcfg = nb.default_config(G=1.0, eps=1e-3, theta=0.5, dt=1e-3, leaf_capacity=32, order=p,
                                 m2l=nb.M2L_PLAIN, p2p=nb.P2P_LEAF)
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
