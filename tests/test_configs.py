"""Workload generators (host logic, no GPU)."""
import numpy as np


def test_c4_ensemble_draws(configs):
    cfgs, steps = configs.c4()
    again, _ = configs.c4()
    assert steps == 500 and len(cfgs) == 64
    edges = {c["nx"] for c in cfgs}
    assert edges <= set(configs.C4_EDGES) and len(edges) >= 4
    cells = sum(c["nx"] ** 3 for c in cfgs)
    assert 2e7 < cells < 1.3e8  # about 4.3e7 expected
    for c, d in zip(cfgs, again):
        assert c["nx"] == c["ny"] == c["nz"]
        assert 0.55 <= c["tau"] <= 1.0
        s = c["sheets"][0]
        assert 16 <= s.rows <= 64 and s.rows == s.cols and s.rows <= c["nx"] - 8
        assert 1e-4 <= s.ks <= 1e-2 and s.ks == s.kb
        # support (floor(X - 2) + 1 .. + 3) inside the grid and off the z walls
        assert np.floor(s.xyz.min() - 2) + 1 >= 2 and np.floor(s.xyz.max() - 2) + 4 <= c["nx"] - 3
        assert np.array_equal(s.xyz, d["sheets"][0].xyz) and c["tau"] == d["tau"]  # seeded
