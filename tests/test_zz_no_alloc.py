"""nbh_direct_fp64_noalloc (the entry point that allocates no device memory)
against nbh_direct_fp64, through the C ABI: identical results on two calls in
a row, the second with fewer targets (the scratch keeps the first call's
contents), and a clear error above its fixed maximum."""
import numpy as np
import pytest

import nbody_amd as nb
from helpers import make_sim

pytestmark = pytest.mark.gpu


def test_noalloc_direct_matches_allocating_entry_point():
    print("[case] noalloc_direct_plummer6000", flush=True)
    n = 6000
    x, v, m = nb.ics.plummer(n, 31)
    sim = make_sim(x, v, m, eps=1e-3, leaf_capacity=16)
    sim.step(1)  # particles now in Morton order: the slot lookup is not the identity
    rng = np.random.default_rng(2)
    for nt in (4000, 37):
        tg = rng.choice(n, nt, replace=False)
        a0, p0 = sim.direct_fp64(tg)
        a1, p1 = sim.direct_fp64_noalloc(tg)
        assert np.array_equal(a0, a1) and np.array_equal(p0, p1), nt
        assert np.all(np.isfinite(a1)) and np.all(np.isfinite(p1))
    with pytest.raises(RuntimeError, match="exceed its fixed scratch"):
        sim.direct_fp64_noalloc(np.zeros(nb._lib.NOALLOC_MAX_TARGETS + 1, dtype=np.int64))
    with pytest.raises(RuntimeError, match="out of range"):
        sim.direct_fp64_noalloc(np.array([0, n]))
    a2, _ = sim.direct_fp64_noalloc(np.array([5, 5, 0]))
    a3, _ = sim.direct_fp64(np.array([5, 5, 0]))
    assert np.array_equal(a2, a3)
    sim.close()
