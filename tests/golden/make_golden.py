"""Writes the seeded golden fixtures of this directory from the FP64 oracle (run from the repository root:
``python tests/golden/make_golden.py``). The reference ships no golden vectors (its tests draw unseeded random
scenes), so these pin the oracle's own outputs on the 1,000-ellipsoid, 64x64 configuration and on the reference's
two-ray gradient scenarios (tests/grad_check.py:31-99)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ever_amd import scenes  # noqa: E402
from oracle import oracle  # noqa: E402

KEYS = ("means", "scales", "quats", "densities", "features")


def config1_fixture():
    sc, o, d = scenes.config1()
    args = [sc[k].double().numpy() for k in KEYS]
    f = oracle.forward(*args, o.double().numpy(), d.double().numpy(), 0.0, 100.0, max_iters=500)
    g = scenes.upstream_grad(o.shape[0]).double().numpy()
    b = oracle.backward(*args, o.double().numpy(), d.double().numpy(), 0.0, 500, g, f)
    E = int(f["iters"].max())
    np.savez_compressed(os.path.join(HERE, "config1_fp64.npz"), color=f["color"], iters=f["iters"],
                        tri_collection=f["tri_collection"].reshape(500, -1)[:E], touch_count=f["touch_count"],
                        grad_out=g, **{k: v for k, v in b.items() if v is not None and k != "dinitial_drgb"})


def gradcheck_fixture():
    # rays (0,0,0) and (0,0,1) along +z, one primitive with quaternion (0,0,0,1), tmin 0.3 (grad_check.py:40-61)
    rng = np.random.default_rng(0)
    out = {}
    for case in range(4):
        means = 2 * rng.random((1, 3)) - 1
        means[0, :2] *= 0.3
        scales = 0.5 * rng.random((1, 3)) + 0.05
        quats = np.array([[0.0, 0.0, 0.0, 1.0]])
        dens = 1.1 * rng.random(1)
        feats = rng.random((1, 1, 3))
        rayo = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.0]]) if case % 2 == 0 else np.array([[0.0, 0.0, 0.0],
                                                                                             [0.0, 0.0, 0.1]])
        rayd = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, 1.0]])
        g = np.array([[1.0, 1, 1, 1, 0], [1.0, 1, 1, 1, 0]]) if case % 2 == 0 else np.array([[1.0, 1, 1, 1, 10]] * 2)
        f = oracle.forward(means, scales, quats, dens, feats, rayo, rayd, 0.3, 100.0)
        b = oracle.backward(means, scales, quats, dens, feats, rayo, rayd, 0.3, 500, g, f)
        for k, v in dict(means=means, scales=scales, quats=quats, densities=dens, features=feats, rayo=rayo,
                         rayd=rayd, grad_out=g, color=f["color"], iters=f["iters"]).items():
            out[f"{case}_{k}"] = v
        for k in ("dmeans", "dscales", "dquats", "ddensities", "dfeatures", "drayo", "drayd"):
            out[f"{case}_{k}"] = b[k]
    np.savez_compressed(os.path.join(HERE, "gradcheck_fp64.npz"), **out)


if __name__ == "__main__":
    config1_fixture()
    gradcheck_fixture()
    print("fixtures written to", HERE)
