"""The measurement script's host logic, without a GPU: the one-GPU guard, the
timing modes time_steps asks the library for, and the names of the step's
same-level kernel that its roofline entry looks up in the profiles."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


class _StubSim:
    """Counts steps and records the timing modes it is given."""

    def __init__(self):
        self.steps = 0
        self.modes = []

    def step(self, n):
        self.steps += n

    def set_timing(self, on=True, serial=False):
        self.modes.append((bool(on), bool(serial)))

    def stats(self):
        keys = ["ms_tree", "ms_moments", "ms_walk", "ms_m2l", "ms_l2l", "ms_p2p", "ms_total", "p2p_interactions",
                "m2l_pairs", "split_pairs", "p2p_pairs", "ms_m2l_same", "m2l_same_pairs", "m2l_same_launches"]
        return {k: 1.0 for k in keys}


def test_time_steps_modes_and_sums():
    bench = _bench()
    sim = _StubSim()
    tot = {}
    agg, wall, inter, m2l, split, p2p, _ = bench.time_steps(sim, 2, 5, serial=False, totals=tot, timing=False)
    # warm-up steps untimed, then the timed ones; events off in the timed steps, and off again after
    assert sim.steps == 7
    assert sim.modes == [(False, False), (False, False)]
    assert agg["ms_total"] == 5.0 and inter == 5 and m2l == 5 and tot["m2l_same_pairs"] == 5
    sim = _StubSim()
    bench.time_steps(sim, 0, 3, serial=True, totals={}, timing=True)
    assert sim.modes == [(True, True), (False, False)] and sim.steps == 3


def test_same_level_kernel_layouts():
    bench = _bench()
    # theta = 1/2, q = 5 with every degree-5 term: 273 rows + the zero row, 64 KiB with the map:
    # two tables per CU (256 threads); theta = 0.4: 116 KiB, one table (512 threads)
    assert bench.same_level_kernel(4, 21, 1, 2) == "m2l_lds_kernel<4, 21, true, 256, true, false>"
    assert bench.same_level_kernel(4, 21, 2, 5) == "m2l_lds_kernel<4, 21, true, 512, true, false>"
    # a table larger than the LDS stays in global memory
    assert bench.same_level_kernel(4, 21, 1, 2, lds_per_cu=32 * 1024) == \
        "m2l_lds_kernel<4, 21, false, 256, true, false>"
