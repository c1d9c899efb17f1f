"""Forward parity of the HIP tracer with the FP64 oracle: colour and opacity within 1e-4 (the reference tests'
tolerance, tests/compare_quadrature.py:65), identical event lists, counts and edge cases."""
import math
import os

import numpy as np
import pytest
import torch

from ever_amd import scenes
from ever_amd import tracer as T
from gpu_util import KEYS, gpu_forward, oracle_forward
from oracle import oracle

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("traversal", ["ray", "wave"])
def test_config1_matches_golden_oracle(traversal):
    print(f"case: config1 forward vs golden, {traversal} traversal", flush=True)
    z = np.load(os.path.join(GOLD, "config1_fp64.npz"))
    sc, o, d = scenes.config1()
    _, out = gpu_forward(sc, o, d, traversal=traversal)
    col = _np(out["color"])
    np.testing.assert_allclose(col[:, :4], z["color"][:, :4], atol=1e-4, rtol=1e-4)
    np.testing.assert_allclose(col[:, 4], z["color"][:, 4], atol=1e-4, rtol=1e-4)
    np.testing.assert_array_equal(_np(out["iters"]), z["iters"])
    E = z["tri_collection"].shape[0]
    tri = _np(out["tri_collection"]).reshape(500, -1)[:E]
    mask = np.arange(E)[:, None] < z["iters"][None, :]
    np.testing.assert_array_equal(tri[mask], z["tri_collection"][mask])
    np.testing.assert_array_equal(_np(out["touch_count"]), z["touch_count"])


@pytest.mark.parametrize("kbuf", [8, 16, 19, 32])
def test_kbuffer_sizes_agree(kbuf):
    print(f"case: kbuf {kbuf}", flush=True)
    sc, o, d = scenes.config1()
    _, a = gpu_forward(sc, o, d, kbuf=16)
    _, b = gpu_forward(sc, o, d, kbuf=kbuf)
    np.testing.assert_array_equal(_np(a["iters"]), _np(b["iters"]))
    np.testing.assert_allclose(_np(a["color"]), _np(b["color"]), rtol=0, atol=1e-6)


def test_deterministic_bitwise():
    sc, o, d = scenes.config1()
    _, a = gpu_forward(sc, o, d)
    _, b = gpu_forward(sc, o, d)
    assert torch.equal(a["color"], b["color"]) and torch.equal(a["iters"], b["iters"])
    assert torch.equal(a["touch_count"], b["touch_count"])


@pytest.mark.parametrize("N", [1, 5, 10, 20, 40])
@pytest.mark.parametrize("inside", [False, True])
def test_quadrature_scenarios(N, inside):
    # tests/compare_quadrature.py:30-102 (seeded): one ray from the origin along +z; outside and inside cases
    print(f"case: quadrature N={N} inside={inside}", flush=True)
    for dm in (0.01, 0.1, 1.0):
        rng = np.random.default_rng(31 * N + int(inside) + int(100 * dm))
        scale = 0.5 * rng.random((N, 3)) + 1e-3
        if inside:
            mean = 1.2 * rng.random((N, 3)) - 0.2
            mean[:, :2] *= 0.5
            tmin = float(rng.random() * 0.3)
        else:
            mean = rng.random((N, 3))
            mean[:, :2] *= 0.5
            mean[:, 2] += 0.5
            tmin = 0.0
        q = 2 * rng.random((N, 4)) - 1
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        sc = dict(means=torch.tensor(mean, dtype=torch.float32), scales=torch.tensor(scale, dtype=torch.float32),
                  quats=torch.tensor(q, dtype=torch.float32),
                  densities=torch.tensor(dm * rng.random(N), dtype=torch.float32),
                  features=torch.tensor(rng.random((N, 1, 3)), dtype=torch.float32))
        o = torch.zeros((1, 3))
        d = torch.tensor([[0.0, 0.0, 1.0]])
        _, out = gpu_forward(sc, o, d, tmin=tmin)
        ref = oracle_forward(sc, o, d, tmin=tmin)
        np.testing.assert_allclose(_np(out["color"])[:, :4], ref["color"][:, :4], atol=1e-4, rtol=1e-4)


def test_start_inside_multi_origin_tmin():
    # rays from several origins inside the scene with tmin > 0 (start-inside density, ray 0's prefilter)
    sc = scenes.uniform_scene(300, 0, seed=4)
    rng = np.random.default_rng(4)
    o = torch.tensor(np.tile(sc["means"][0].numpy(), (256, 1)) + rng.normal(0, 1e-3, (256, 3)), dtype=torch.float32)
    d = torch.tensor(rng.normal(size=(256, 3)), dtype=torch.float32)
    _, out = gpu_forward(sc, o, d, tmin=0.01)
    ref = oracle_forward(sc, o, d, tmin=0.01)
    assert len(ref["initial_inds"]) > 0 and (ref["initial_drgb"][:, 0] > 0).any()
    np.testing.assert_allclose(_np(out["initial_drgb"]), ref["initial_drgb"], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(_np(out["color"])[:, :4], ref["color"][:, :4], atol=1e-4, rtol=1e-4)


def test_max_iters_and_tmax():
    sc, o, d = scenes.config1()
    for mi, tmax in ((3, 100.0), (500, 3.0), (0, 100.0)):
        _, out = gpu_forward(sc, o, d, max_iters=mi, tmax=tmax)
        ref = oracle_forward(sc, o, d, max_iters=mi, tmax=tmax)
        np.testing.assert_array_equal(_np(out["iters"]), ref["iters"])
        np.testing.assert_allclose(_np(out["color"])[:, :4], ref["color"][:, :4], atol=1e-4, rtol=1e-4)


def test_single_primitive_and_empty_inputs():
    sc = scenes.uniform_scene(1, 0, seed=2)
    sc["means"][:] = torch.tensor([0.0, 0.0, 0.0])
    o, d = scenes.pinhole_rays(8, 8, (0.0, 0.0, -3.0))
    _, out = gpu_forward(sc, o, d)
    ref = oracle_forward(sc, o, d)
    np.testing.assert_allclose(_np(out["color"])[:, :4], ref["color"][:, :4], atol=1e-4, rtol=1e-4)
    # no rays
    _, out = gpu_forward(sc, o[:0], d[:0])
    assert out["color"].shape == (0, 5)
    # no primitives: every ray is empty
    sc0 = {k: v[:0] for k, v in scenes.uniform_scene(4, 0, seed=2).items()}
    _, out = gpu_forward(sc0, o, d)
    assert float(out["color"].abs().max()) == 0.0 and int(out["iters"].max()) == 0


def test_clustered_needles_subset_matches_oracle():
    # 100k clustered anisotropic ellipsoids (aspect up to 100, densities up to 200), SH3 fused colour
    print("case: clustered 100k", flush=True)
    sc = scenes.clustered_scene(100_000, 3, seed=5)
    o, d = scenes.pinhole_rays(320, 180, (0.0, 0.0, -4.0))
    eye = (0.0, 0.0, -4.0)
    _, out = gpu_forward(sc, o, d, color_mode=1, sh_origin=eye, sh_degree=3, activation=1)
    idx = np.arange(0, o.shape[0], 29)
    rgb = oracle.eval_sh(sc["means"].double().numpy(), sc["features"].double().numpy(), 3, np.array(eye))
    f0 = ((np.logaddexp(0.0, 10 * rgb) / 10 - 0.5) / 0.28209479177387814).reshape(-1, 1, 3)
    args = [sc["means"].double().numpy(), sc["scales"].double().numpy(), sc["quats"].double().numpy(),
            sc["densities"].double().numpy(), f0]
    ref = oracle.forward(*args, o.double().numpy()[idx], d.double().numpy()[idx], 0.0, 100.0, use_bvh=True)
    np.testing.assert_allclose(_np(out["color"])[idx, :4], ref["color"][:, :4], atol=1e-4, rtol=1e-4)
    assert (_np(out["iters"])[idx] == ref["iters"]).mean() > 0.999


@pytest.mark.parametrize("traversal", ["ray", "wave"])
def test_events_by_slot_are_the_by_ray_events_permuted(traversal):
    # a batch large enough to be sorted (processing order != ray order); by-slot lists are the by-ray lists with
    # columns permuted by ray_order, every other output is unchanged and so is the adjoint that reads them
    print(f"case: events by slot vs by ray, {traversal}", flush=True)
    sc = scenes.clustered_scene(20_000, 3, seed=21)
    o, d = scenes.pinhole_rays(320, 240, (0.0, 0.0, -4.0))
    D = torch.device("cuda:0")
    t = [sc[k].to(D).contiguous() for k in KEYS]
    scene = T.Scene(*t, color_mode=1, sh_origin=(0.0, 0.0, -4.0), sh_degree=3, activation=1).build()
    og, dg = o.to(D).contiguous(), d.to(D).contiguous()
    a = T.trace_forward(scene, og, dg, 0.0, 100.0, 500, traversal=traversal)
    b = T.trace_forward(scene, og, dg, 0.0, 100.0, 500, traversal=traversal, event_layout="slot",
                        count_touches=False)
    nr = o.shape[0]
    order = a["ray_order"].long()
    assert torch.equal(order, b["ray_order"].long())
    assert not torch.equal(order, torch.arange(nr, device=D))
    for k in ("color", "iters", "states", "ray_state64", "initial_drgb"):
        assert torch.equal(a[k], b[k]), k
    assert b["touch_count"] is None
    ta, tb = a["tri_collection"].view(500, nr), b["tri_collection"].view(500, nr)
    mask = torch.arange(500, device=D)[:, None] < a["iters"][order][None, :]
    assert torch.equal(ta[:, order][mask], tb[mask])
    g = scenes.upstream_grad(nr, seed=5).to(D)
    ga = T.trace_backward(scene, og, dg, 0.0, 100.0, 500, a, g)
    gb = T.trace_backward(scene, og, dg, 0.0, 100.0, 500, b, g)
    for k in ga:  # per-primitive sums are float atomics: equal up to summation order
        assert float((ga[k] - gb[k]).norm()) <= 1e-6 * float(ga[k].norm()) + 1e-12, k


@pytest.mark.parametrize("traversal", ["ray", "wave"])
def test_last_dirac_is_the_last_integrated_control_point(traversal):
    # SavedForBackward.diracs (shaders.slang:229-232): +-density * (1, SH_C0 * f0 + 0.5) of each ray's last
    # integrated event (entry +, exit -), zeros for a ray without events
    sc, o, d = scenes.config1()
    D = torch.device("cuda:0")
    t = [sc[k].to(D).contiguous() for k in KEYS]
    scene = T.Scene(*t).build()
    out = T.trace_forward(scene, o.to(D).contiguous(), d.to(D).contiguous(), 0.0, 100.0, 500, traversal=traversal,
                          last_dirac=True)
    it = out["iters"].cpu().numpy()
    tri = out["tri_collection"].cpu().numpy().reshape(500, -1)
    nr = o.shape[0]
    dens = sc["densities"].numpy().astype(np.float64)
    col = sc["features"][:, 0, :].numpy().astype(np.float64) * 0.28209479177387814 + 0.5
    want = np.zeros((nr, 4))
    for r in np.nonzero(it > 0)[0]:
        ev = tri[it[r] - 1, r]
        h = dens[ev >> 1] * (1.0 if ev & 1 else -1.0)
        want[r] = [h, h * col[ev >> 1, 0], h * col[ev >> 1, 1], h * col[ev >> 1, 2]]
    assert (it == 0).any() and (it > 0).any()
    np.testing.assert_allclose(out["last_dirac"].cpu().numpy(), want, rtol=1e-6, atol=1e-6)


def test_three_million_primitives_match_oracle():
    # a large scene: 12M leaf slots in the sort and the tree (int32 indices throughout), 53 events per ray
    print("case: 3M small primitives, 48x48", flush=True)
    sc = scenes.uniform_scene(3_000_000, 0, seed=41)
    sc["scales"] = (sc["scales"] * 0.2).contiguous()
    o, d = scenes.pinhole_rays(48, 48, (0.0, 0.0, -3.0))
    ref = oracle_forward(sc, o, d, use_bvh=True)
    for traversal in ("ray", "wave"):
        _, out = gpu_forward(sc, o, d, traversal=traversal)
        assert (_np(out["iters"]) != ref["iters"]).mean() <= 1e-3
        np.testing.assert_allclose(_np(out["color"])[:, :4], ref["color"][:, :4], atol=1e-4, rtol=1e-4)
