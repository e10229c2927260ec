"""CPU tests of the C ABI boundary (no GPU needed): the library loads, exports
every symbol include/nbody_hip.h declares, validates configurations before
touching a device, reports errors through nbh_last_error, and its seeded
initial conditions match the Plummer model's closed forms."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import nbody_amd as nb
from nbody_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nbody_hip.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nbh_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = nb.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)


def test_exported_symbols_via_nm():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (nbh_[a-z0-9_]+)", out))
    assert set(declared_symbols()) <= exported


def test_abi_and_config_layout():
    L = nb.lib()
    assert L.nbh_abi_version() == 3
    assert L.nbh_config_size() == ctypes.sizeof(nb.Config)


def test_config_defaults_are_the_reference_constants():
    c = nb.default_config()
    assert c.G == 1.0          # FORCE_CONSTANT = -1 (field.cl:7-9)
    assert c.eps == 0.01       # PARTICLE_RADIUS (field.cl:3-5)
    assert c.theta == 0.5      # NODE_APPROX_RATIO (interaction.cl:3-5)
    assert c.leaf_capacity == 8  # open_cl_simulation.cpp:41-47
    assert c.solver == nb.SOLVER_FMM and c.integrator == nb.INT_KDK


def _create(cfg, n=16):
    x = np.random.default_rng(0).random((n, 3))
    v = np.zeros((n, 3))
    m = np.full(n, 1.0 / n)
    h = ctypes.c_void_p()
    dp = ctypes.POINTER(ctypes.c_double)
    rc = nb.lib().nbh_create(ctypes.byref(cfg), n, x.ctypes.data_as(dp), v.ctypes.data_as(dp),
                             m.ctypes.data_as(dp), None, ctypes.byref(h))
    return rc, nb.lib().nbh_last_error(None).decode(), h


@pytest.mark.parametrize("field,value,msg", [
    ("theta", 0.333333, "theta"),
    ("theta", 0.0, "theta"),
    ("order", 0, "order"),
    ("order", 9, "order"),
    ("leaf_capacity", 0, "leaf_capacity"),
    ("eps", -1.0, "eps"),
    ("dt", 0.0, "dt"),
    ("max_depth", 22, "max_depth"),
    ("integrator", 7, "integrator"),
    ("force_law", nb.LAW_COMPAT_NAIVE, "direct solver"),
    ("struct_size", 8, "struct_size"),
])
def test_invalid_configs_rejected_before_device(field, value, msg):
    cfg = nb.default_config()
    setattr(cfg, field, value)
    rc, err, h = _create(cfg)
    assert rc == _lib.NBH_E_INVALID
    assert msg in err
    assert not h.value


def test_nonfinite_input_rejected():
    cfg = nb.default_config()
    n = 8
    x = np.zeros((n, 3)); x[3, 1] = np.nan
    h = ctypes.c_void_p()
    dp = ctypes.POINTER(ctypes.c_double)
    m = np.ones(n)
    rc = nb.lib().nbh_create(ctypes.byref(cfg), n, x.ctypes.data_as(dp), x.ctypes.data_as(dp),
                             m.ctypes.data_as(dp), None, ctypes.byref(h))
    assert rc == _lib.NBH_E_NONFINITE


def test_no_device_is_a_clean_error():
    if nb.device_count() > 0:
        pytest.skip("a GPU is visible")
    rc, err, h = _create(nb.default_config())
    assert rc == _lib.NBH_E_HIP and "no HIP device" in err
    with pytest.raises(RuntimeError, match="no HIP device"):
        x, v, m = nb.ics.uniform(10, 0)
        nb.HipSimulation.from_arrays(x, v, m, config=nb.default_config())


def test_theta_rationals_accepted():
    # validation accepts num/den with den <= 64; create then fails only for lack of a device
    for th in (0.5, 0.4, 0.75, 0.25):
        cfg = nb.default_config(theta=th)
        rc, err, h = _create(cfg)
        assert rc in (_lib.NBH_OK, _lib.NBH_E_HIP), err
        if rc == _lib.NBH_OK:
            nb.lib().nbh_destroy(h)


# ---------------------------------------------------------------------------
# seeded initial conditions
def test_uniform_ics():
    x, v, m = nb.ics.uniform(20000, 0)
    assert x.min() >= 0 and x.max() < 1 and np.all(v == 0) and np.allclose(m, 1 / 20000)
    assert np.allclose(x.mean(0), 0.5, atol=0.01)
    x2, _, _ = nb.ics.uniform(20000, 0)
    assert np.array_equal(x, x2)
    x3, _, _ = nb.ics.uniform(20000, 1)
    assert not np.array_equal(x, x3)


def test_plummer_ics_match_closed_forms():
    n = 200000
    x, v, m = nb.ics.plummer(n, 1, rmax=20.0)
    r = np.linalg.norm(x, axis=1)
    assert r.max() <= 20.0 and np.isclose(m.sum(), 1.0)
    # cumulative mass M(<r) = r^3 / (1 + r^2)^1.5, renormalised to the truncation
    mt = 20.0 ** 3 / (1 + 400.0) ** 1.5
    for rr in (0.5, 1.0, 2.0, 5.0):
        frac = (r < rr).mean()
        assert abs(frac - rr ** 3 / (1 + rr ** 2) ** 1.5 / mt) < 0.005
    # bound: speeds below the local escape speed sqrt(2) (1 + r^2)^(-1/4)
    s = np.linalg.norm(v, axis=1)
    assert np.all(s <= np.sqrt(2) * (1 + r ** 2) ** -0.25 + 1e-12)
    # isotropy and the untruncated kinetic energy 3 pi / 64 (truncation at 20a changes it < 1%)
    assert abs(v.mean(0)).max() < 0.01
    assert abs(0.5 * (m * s ** 2).sum() - 3 * np.pi / 64) < 0.01 * 3 * np.pi / 64


def test_ensemble_ics():
    x, v, m, off = nb.ics.ensemble(nsys=8, nmin=512, nmax=4096)
    assert off[0] == 0 and off[-1] == len(x) and np.all(np.diff(off) >= 512)
    for s in range(8):
        assert np.isclose(m[off[s]:off[s + 1]].sum(), 1.0)


def test_cli_fails_cleanly_without_device():
    exe = os.path.join(os.path.dirname(_lib.LIB_PATH), "nbody_cli")
    assert os.path.exists(exe)
    if nb.device_count() > 0:
        pytest.skip("a GPU is visible")
    r = subprocess.run([exe, "--n", "10"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "no HIP device" in r.stderr
