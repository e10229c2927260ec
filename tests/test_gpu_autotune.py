"""nbh_autotune through the C ABI: the leaf capacity and P2P tile chosen from
the candidates, set on the handle, the state left unchanged, and the forces
afterwards those of a fresh handle built with the chosen configuration."""
import numpy as np
import pytest

import nbody_amd as nb
from helpers import make_sim, rms_rel

pytestmark = pytest.mark.gpu


def test_autotune_picks_a_candidate_and_keeps_the_state():
    print("[case] autotune_plummer32768", flush=True)
    n = 32768
    x, v, m = nb.ics.plummer(n, 7)
    sim = make_sim(x, v, m, leaf_capacity=8, order=4, eps=1e-3, dt=1e-3)
    sim.step(2)
    x0, v0, _, _, t0 = sim.state()
    cap, chunk, ms, table = sim.autotune([16, 32, 64], [64, 128], evals=2)
    print(f"  chose cap {cap} chunk {chunk}: {ms:.3f} ms/eval; table {table}", flush=True)
    assert (cap, chunk) in table and cap in (16, 32, 64) and chunk in (64, 128)
    assert ms == min(table.values()) and all(t > 0 for t in table.values())
    assert sim.config.leaf_capacity == cap and sim.config.p2p_chunk == chunk
    x1, v1, _, _, t1 = sim.state()
    assert np.array_equal(x0, x1) and np.array_equal(v0, v1) and t0 == t1
    a_tuned, _ = sim.forces()
    ref = make_sim(x1, v1, m, leaf_capacity=cap, p2p_chunk=chunk, order=4, eps=1e-3, dt=1e-3)
    a_ref, _ = ref.forces()
    e = rms_rel(a_tuned, a_ref)
    print(f"  forces vs a fresh handle with the chosen configuration: rms {e:.2e}", flush=True)
    assert e < 1e-6
    assert sim.debug_sizes()["leaves"] == ref.debug_sizes()["leaves"]
    # the chosen configuration steps on
    sim.step(1)
    ref.step(1)
    assert rms_rel(sim.state()[0], ref.state()[0]) < 1e-10
    ref.close()
    sim.close()


def test_autotune_rejects_bad_candidates():
    print("[case] autotune_invalid", flush=True)
    x, v, m = nb.ics.plummer(2048, 3)
    sim = make_sim(x, v, m, leaf_capacity=16)
    with pytest.raises(RuntimeError, match="chunk"):
        sim.autotune([16], [96])
    with pytest.raises(RuntimeError, match="capacity"):
        sim.autotune([0], None)
    assert sim.config.leaf_capacity == 16
    cap, chunk, ms, table = sim.autotune(None, None, evals=1)
    assert (cap, chunk) == (16, 64) and list(table) == [(None, None)]
    sim.close()
    d = make_sim(x, v, m, solver=nb.SOLVER_DIRECT)
    with pytest.raises(RuntimeError, match="FMM"):
        d.autotune([16], None)
    d.close()
