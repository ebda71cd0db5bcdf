"""The oracle still reproduces the committed golden fixtures
(tests/golden/golden_cfg2.npz, written by tests/golden/make_golden.py)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))


def test_oracle_reproduces_golden():
    import make_golden
    g = np.load(os.path.join(HERE, "golden", "golden_cfg2.npz"))
    r = make_golden.compute()
    assert np.array_equal(r["idx"], g["idx"])
    assert (g["idx"] >= 0).sum() > 1000
    for key in ("pts", "nrm", "sd", "albedo"):
        assert np.array_equal(r[key], g[key]), key
    for key in ("w", "dw", "gin"):
        assert np.array_equal(r[key].astype(np.float32), g[key]), key
    for key in ("img", "gp", "gn", "gs", "ga"):
        assert np.allclose(r[key], g[key], rtol=1e-12, atol=1e-14), key
