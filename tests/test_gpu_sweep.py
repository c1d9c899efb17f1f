"""Randomised parity sweep of the HIP tracer against the FP64 oracle: seeded draws of scene kind and size,
feature degree, camera (outside or inside the scene), tmin, tmax, max_iters, k-buffer size and tree walk, each
checked forward (colour and opacity 1e-4 absolute, distortion 1e-4 relative, per-ray event counts exact) and
backward (every gradient tensor: 1e-3 relative L2 error, and the largest single error 1e-3 of the tensor's
largest element). The same algorithm evaluated in float32 on the CPU is printed beside each figure: it shows how
far the case is conditioned at float32 precision (it misses these bars on scale gradients along the thin axis of
needle-like primitives, which the tracer's FP64 root adjoint holds).
Cases with a tmax that cuts through primitives and k-buffer sizes other than 16 check that results follow the
reference's 16-event rounds whatever the buffer size. The scenes are traced in DC colour mode, where features of
degree 3 colour events by their DC term and the backward is that forward's exact adjoint (both here and in the
oracle): such cases exercise the degree-0 path with (N,16,3) inputs, not the SH evaluation (which has its own tests
and the full-size configuration tests through trace_rays_sh)."""
import math
import os

import numpy as np
import pytest
import torch

from ever_amd import scenes
from ever_amd import tracer as T
from gpu_util import KEYS
from oracle import oracle

pytestmark = pytest.mark.gpu
D = torch.device("cuda:0")
GRADS = ("dmeans", "dscales", "dquats", "ddensities", "dfeatures", "drayo", "drayd")


def _draw(seed):
    rng = np.random.default_rng(1000 + seed)
    kind = ["uniform", "clustered"][seed % 2]
    deg = int(rng.choice([0, 3]))
    if kind == "uniform":
        sc = scenes.uniform_scene(int(rng.integers(100, 3000)), deg, seed=seed)
    else:
        sc = scenes.clustered_scene(int(rng.integers(2000, 20000)), deg, seed=seed)
    inside = rng.random() < 0.3
    if inside:
        eye = tuple(float(v) for v in sc["means"][int(rng.integers(0, sc["means"].shape[0]))])
        tmin = 0.05
    else:
        v = rng.normal(size=3)
        dist = rng.uniform(2.5, 4.0) if kind == "uniform" else rng.uniform(3.5, 5.0)
        eye = tuple(float(x) for x in v / np.linalg.norm(v) * dist)
        tmin = float(rng.choice([0.0, 0.1]))
    o, d = scenes.pinhole_rays(48, 36, eye, fov_deg=float(rng.uniform(25, 60)),
                               target=tuple(float(x) for x in rng.normal(0, 0.2, 3)))
    # tmax: unbounded, or cutting through the scene
    tmax = 100.0 if rng.random() < 0.6 else float(np.linalg.norm(eye)) + float(rng.uniform(-0.5, 1.0))
    max_iters = 500 if rng.random() < 0.6 else int(rng.integers(5, 60))
    kbuf = int(rng.choice([8, 16, 19, 32]))
    traversal = str(rng.choice(["ray", "wave"]))
    if rng.random() < 0.3:
        # directions of any length: the start point moves by tmin times the raw direction (shaders.slang:178-179)
        d = (d * torch.from_numpy(rng.uniform(0.5, 2.0, (d.shape[0], 1)).astype(np.float32))).contiguous()
    wcts = None
    if rng.random() < 0.4:
        # the means2D statistic: the camera's world-to-pixel matrix (ever.py:147-163 layout)
        W = scenes.world_to_camera(eye, tuple(float(x) for x in rng.normal(0, 0.2, 3))).numpy()
        K = np.eye(4)
        K[0, 0] = K[1, 1] = 40.0
        wcts = (W @ K).astype(np.float32)[None]
    g = torch.from_numpy(rng.normal(size=(o.shape[0], 5)).astype(np.float32))
    # the tracer takes tmin and tmax in float32: the oracle gets the same values (a start point within 1e-8 of a
    # surface is placed differently by 0.05 and its float32 rounding)
    tmin, tmax = float(np.float32(tmin)), float(np.float32(tmax))
    return dict(sc=sc, o=o, d=d, tmin=tmin, tmax=tmax, max_iters=max_iters, kbuf=kbuf, traversal=traversal, g=g,
                wcts=wcts,
                desc=f"seed {seed}: {kind} N={sc['means'].shape[0]} deg {deg} inside {inside} tmin {tmin} "
                     f"tmax {tmax:.2f} max_iters {max_iters} K {kbuf} {traversal} wcts {wcts is not None}")


# seeds beyond the first 96 that once failed: grazing hits, near-tied events ordered differently by float32 and FP64
# depths, exits beyond a cutting tmax at the reference's round boundaries, thin-axis scale gradients, a camera inside
# a primitive whose surface passes within 1e-8 of the start point (float32 and FP64 roots, or containment tests, on
# opposite sides of it)
HARD_SEEDS = (138, 156, 180, 204, 289, 311, 330, 342, 346, 348, 354, 369, 438, 468, 494, 496, 513, 522, 546, 1483,
              2059)


def _seeds():
    # the committed seeds; EVER_SWEEP_SEEDS="a-b" draws seeds a..b instead (wider searches)
    spec = os.environ.get("EVER_SWEEP_SEEDS")
    if not spec:
        return list(range(96)) + list(HARD_SEEDS)
    a, b = (int(x) for x in spec.split("-"))
    return range(a, b + 1)


@pytest.mark.parametrize("seed", _seeds())
def test_random_configuration_matches_oracle(seed):
    c = _draw(seed)
    print("case:", c["desc"], flush=True)
    sc, o, d = c["sc"], c["o"], c["d"]
    t = [sc[k].to(D).contiguous() for k in KEYS]
    og, dg = o.to(D).contiguous(), d.to(D).contiguous()
    scene = T.Scene(*t).build()
    f = T.trace_forward(scene, og, dg, c["tmin"], c["tmax"], c["max_iters"], kbuf=c["kbuf"],
                        traversal=c["traversal"])
    wg = None if c["wcts"] is None else torch.from_numpy(c["wcts"]).to(D)
    b = T.trace_backward(scene, og, dg, c["tmin"], c["tmax"], c["max_iters"], f, c["g"].to(D), wcts=wg,
                         kbuf=c["kbuf"])
    torch.cuda.synchronize()
    npsc = [sc[k].double().numpy() for k in KEYS]
    ref = oracle.forward(*npsc, o.double().numpy(), d.double().numpy(), c["tmin"], c["tmax"],
                         max_iters=c["max_iters"], use_bvh=True)
    if seed % 3 == 0:
        # a proposal bounding each ray's first traversal (any depths, including ones before a cutting tmax)
        # leaves the result bit-identical to the plain trace
        prop = torch.from_numpy(np.random.default_rng(seed).uniform(0.2, 6.0, o.shape[0]).astype(np.float32))
        fp = T.trace_forward(scene, og, dg, c["tmin"], c["tmax"], c["max_iters"], kbuf=c["kbuf"],
                             traversal=c["traversal"], proposal=prop.to(D))
        assert torch.equal(fp["color"], f["color"]) and torch.equal(fp["iters"], f["iters"])
    # the same algorithm evaluated in float32 on the CPU: its distance to the FP64 oracle is the conditioning of
    # the case at float32 precision (reported, not used in the bars)
    r32 = oracle.forward(*npsc, o.double().numpy(), d.double().numpy(), c["tmin"], c["tmax"],
                         max_iters=c["max_iters"], use_bvh=True, dtype=np.float32)
    col = f["color"].cpu().double().numpy()
    err = np.abs(col[:, :4] - ref["color"][:, :4]).max()
    f_err = np.abs(r32["color"][:, :4] - ref["color"][:, :4]).max()
    dden = max(np.abs(ref["color"][:, 4]).max(), 1e-6)
    derr = np.abs(col[:, 4] - ref["color"][:, 4]).max() / dden
    f_derr = np.abs(r32["color"][:, 4] - ref["color"][:, 4]).max() / dden
    iters = f["iters"].cpu().numpy()
    mism = float((iters != ref["iters"]).mean())
    print(f"colour max abs {err:.2e} (float32 restatement {f_err:.2e}), distortion rel {derr:.2e} ({f_derr:.2e}), "
          f"iters mismatch {mism:.4f}, events/ray {iters.mean():.1f}", flush=True)
    assert err <= 1e-4 and derr <= 1e-4 and mism == 0.0, (err, derr, mism)
    w64 = None if c["wcts"] is None else c["wcts"].astype(np.float64)
    rb = oracle.backward(*npsc, o.double().numpy(), d.double().numpy(), c["tmin"], c["max_iters"],
                         c["g"].double().numpy(), ref, wcts=w64)
    b32 = oracle.backward(*npsc, o.double().numpy(), d.double().numpy(), c["tmin"], c["max_iters"],
                          c["g"].double().numpy(), r32, wcts=w64, dtype=np.float32)
    for k in GRADS + (("dmeans2D",) if w64 is not None else ()):
        a = b[k].cpu().double().numpy().reshape(rb[k].shape)
        r = rb[k]
        e_l2 = np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-12)
        e_mx = np.abs(a - r).max() / max(np.abs(r).max(), 1e-12)
        r32k = b32[k].astype(np.float64).reshape(r.shape)
        f_l2 = np.linalg.norm(r32k - r) / max(np.linalg.norm(r), 1e-12)
        f_mx = np.abs(r32k - r).max() / max(np.abs(r).max(), 1e-12)
        print(f"  {k}: L2 rel {e_l2:.2e}, max rel {e_mx:.2e} (float32 restatement: {f_l2:.2e}, {f_mx:.2e})",
              flush=True)
        assert e_l2 <= 1e-3 and e_mx <= 1e-3, k
