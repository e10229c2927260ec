"""The multi-core CPU backend (include/nbody/cpu_simulation.h through
libnbody_cpu.so) against the oracle, the CLI's CPU backend, and the CSV
viewer.  No GPU needed."""
import os
import subprocess

import numpy as np
import pytest

import nbody_amd as nb
import pyoracle as po
from nbody_amd.cpu import INT_EULER, INT_KDK, INT_KICK_DRIFT, CpuSimulation
from nbody_amd import viewer


def _ics(n=300, seed=3):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 3))
    v = 0.1 * rng.standard_normal((n, 3))
    m = rng.random(n) / n + 1e-4
    q = m * (1 + 0.5 * rng.random(n))
    return x, v, m, q


def test_forces_match_oracle_direct():
    x, v, m, q = _ics()
    sim = CpuSimulation.from_arrays(x, v, m, q, G=0.7, eps=0.02, threads=4)
    a, p = sim.forces()
    ao, po_ = po.direct(x, q, eps=0.02)
    # oracle (G = 1): acc_i = sum_j q_j r/s^3, pot_i = -sum_j q_j/s; the
    # backend scales both by G and the acceleration by q_i / m_i
    assert np.allclose(a, 0.7 * ao * (q / m)[:, None], rtol=1e-12, atol=1e-14)
    assert np.allclose(p, 0.7 * po_, rtol=1e-12) and np.all(p < 0)
    sim.close()


@pytest.mark.parametrize("integ,mode", [(INT_KDK, 0), (INT_EULER, 1), (INT_KICK_DRIFT, 2)])
def test_integrators_match_oracle(integ, mode):
    x, v, m, _ = _ics(200, 5)
    sim = CpuSimulation.from_arrays(x, v, m, None, time_step=1e-3, G=1.0, eps=1e-2, integrator=integ)
    t = sim.step(7)
    xs, vs, _, _, tt = sim.state()
    xo, vo = po.integrate_direct(x, v, m, m, 1.0, 1e-2, 1e-3, 7, mode)
    assert abs(t - 7e-3) < 1e-15 and tt == t
    assert np.allclose(xs, xo, rtol=1e-12, atol=1e-14) and np.allclose(vs, vo, rtol=1e-12, atol=1e-14)
    sim.close()


def test_dropin_interface_input_order():
    rng = np.random.default_rng(1)
    parts = [nb.Particle(tuple(rng.random(3)), (0.0, 0.0, 0.0), 1e-3, 1e-3) for _ in range(50)]
    sim = CpuSimulation((1.0, 1.0, 1.0, 0.0), parts, 1e-3, None)
    out0 = sim.particles()
    assert np.allclose([p.position for p in out0], [p.position for p in parts])
    t = sim.step()
    assert abs(t - 1e-3) < 1e-15
    out = sim.particles()
    assert len(out) == 50 and out[3].mass == pytest.approx(1e-3)
    sim.close()


def test_errors_are_reported():
    with pytest.raises(RuntimeError):
        CpuSimulation.from_arrays(np.zeros((2, 3)), np.zeros((2, 3)), np.array([1.0, -1.0]))
    with pytest.raises(RuntimeError):
        CpuSimulation.from_arrays(np.zeros((2, 3)), np.zeros((2, 3)), np.ones(2), integrator=7)


def test_cli_cpu_backend_and_viewer(tmp_path):
    exe = os.path.join(os.path.dirname(nb._lib.LIB_PATH), "nbody_cli")
    out = tmp_path / "traj.csv"
    r = subprocess.run([exe, "--backend", "cpu", "--ic", "plummer", "--n", "64", "--dt", "0.001",
                        "--t-end", "0.004", "--csv", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    times, P = viewer.read_csv(str(out))
    assert P.shape == (4, 64, 3) and np.allclose(times, [1e-3, 2e-3, 3e-3, 4e-3], atol=1e-6)
    png = tmp_path / "last.png"
    assert viewer.main([str(out), "--frame", "-1", "--save", str(png)]) == 0
    assert png.stat().st_size > 1000


def test_cli_cpu_fmm_backend_follows_direct(tmp_path):
    """--backend cpu-fmm (CpuSimulation's FMM) against --backend cpu (its
    direct sum) on the same seeded ICs: the trajectories agree to the FMM's
    force error over four steps."""
    exe = os.path.join(os.path.dirname(nb._lib.LIB_PATH), "nbody_cli")
    runs = {}
    for backend in ("cpu", "cpu-fmm"):
        out = tmp_path / f"{backend}.csv"
        r = subprocess.run([exe, "--backend", backend, "--ic", "plummer", "--n", "300", "--dt", "0.001",
                            "--t-end", "0.004", "--order", "6", "--csv", str(out)],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        runs[backend] = viewer.read_csv(str(out))
    (td, Pd), (tf, Pf) = runs["cpu"], runs["cpu-fmm"]
    assert np.allclose(td, tf) and Pd.shape == Pf.shape == (4, 300, 3)
    assert np.max(np.abs(Pf - Pd)) < 1e-5
