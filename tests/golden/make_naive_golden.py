"""Writes golden vectors of the reference NaiveSimulation::step
(src/naive_simulation.cpp:7-46), compiled from the reference's own sources into
oracle/_ref/libnaive_ref.so by oracle/Makefile (`make -C oracle ref`).

Inputs: seeded uniform positions in [0,1)^3 with distinct x coordinates (an x
tie makes the reference divide 0/0, naive_simulation.cpp:10-17), random
velocities, masses and charges; force constant and time step as below.
Outputs: the particles after k steps and the returned time, float32 bit
patterns.  Run from the repository root:  python tests/golden/make_naive_golden.py
"""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402

CASES = [
    # name, n, seed, force_constant, dt, steps
    ("naive_n64_s3", 64, 11, 1.0, 1e-3, 3),
    ("naive_n256_s2", 256, 12, 0.5, 1e-3, 2),
    ("naive_n200_s5_neg", 200, 13, -1.0, 2e-3, 5),
]


def inputs(n, seed):
    rng = np.random.default_rng(seed)
    while True:
        x = rng.random((n, 3)).astype(np.float32)
        if len(np.unique(x[:, 0])) == n:
            break
    v = (rng.random((n, 3)) - 0.5).astype(np.float32) * 0.1
    m = (1.0 + 9.0 * rng.random(n)).astype(np.float32)
    q = (0.1 + 0.9 * rng.random(n)).astype(np.float32)
    return po.particles_to_p12(x, v, m, q)


def main():
    if po.ref_lib() is None:
        raise SystemExit("oracle/_ref/libnaive_ref.so is not built: run `make -C oracle ref`")
    out_dir = os.path.dirname(os.path.abspath(__file__))
    src = "/root/reference/src/naive_simulation.cpp"
    meta = {
        "generator": "tests/golden/make_naive_golden.py",
        "reference_file": "src/naive_simulation.cpp",
        "reference_sha256": hashlib.sha256(open(src, "rb").read()).hexdigest(),
        "compiler": subprocess.run(["g++", "--version"], capture_output=True, text=True).stdout.splitlines()[0],
        "flags": "-std=c++14 -O2 -ffp-contract=off",
        "cases": {},
    }
    for name, n, seed, fc, dt, steps in CASES:
        P = inputs(n, seed)
        out, t = po.naive_reference(P, fc, dt, steps)
        np.savez(os.path.join(out_dir, name + ".npz"), inp=P, out=out, time=np.float32(t),
                 force_constant=np.float32(fc), dt=np.float32(dt), steps=np.int32(steps))
        meta["cases"][name] = {"n": n, "seed": seed, "force_constant": fc, "dt": dt, "steps": steps,
                               "finite": bool(np.all(np.isfinite(out)))}
        print(name, "written")
    with open(os.path.join(out_dir, "naive_golden.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main()
