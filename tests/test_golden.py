"""The oracle reproduces the committed golden fixtures (regression pin of the checker itself)."""
import os

import numpy as np

from ever_amd import scenes
from oracle import oracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KEYS = ("means", "scales", "quats", "densities", "features")


def test_config1_fixture_reproduced():
    z = np.load(os.path.join(HERE, "config1_fp64.npz"))
    sc, o, d = scenes.config1()
    args = [sc[k].double().numpy() for k in KEYS]
    f = oracle.forward(*args, o.double().numpy(), d.double().numpy(), 0.0, 100.0)
    np.testing.assert_allclose(f["color"], z["color"], rtol=1e-12, atol=1e-14)
    np.testing.assert_array_equal(f["iters"], z["iters"])
    E = z["tri_collection"].shape[0]
    np.testing.assert_array_equal(f["tri_collection"].reshape(500, -1)[:E], z["tri_collection"])
    b = oracle.backward(*args, o.double().numpy(), d.double().numpy(), 0.0, 500, z["grad_out"], f)
    for k in ("dmeans", "dscales", "dquats", "ddensities", "dfeatures", "drayo", "drayd"):
        np.testing.assert_allclose(b[k], z[k], rtol=1e-10, atol=1e-12)
