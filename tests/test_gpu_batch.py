"""Several cameras in one batch: a scene built once with per-view SH colours and traced in one launch must give,
ray for ray, the forward results of tracing each camera on its own (bit for bit) and the sum of their gradients;
the persistent launch form must equal the one-wave-per-block form; a HIP graph of build + forward + backward must
reproduce the eager results on every replay; the autotuner must return a measured candidate and reuse its cache."""
import os

import pytest
import torch

from ever_amd import batch as B
from ever_amd import scenes
from ever_amd import tracer as T
from gpu_util import KEYS

pytestmark = pytest.mark.gpu
D = torch.device("cuda:0")
CAMS = [(e, w, h) for e, (w, h) in zip(B.fibonacci_eyes(4, 4.0), [(96, 64), (128, 128), (200, 100), (64, 48)])]


def _scene_tensors(n=30_000, seed=12):
    sc = scenes.uniform_scene(n, 3, seed)
    return [sc[k].to(D).contiguous() for k in KEYS]


def _rel(a, b):
    return float((a - b).norm() / max(float(b.norm()), 1e-12))


@pytest.mark.parametrize("tile, tmax, kbuf", [(0, 100.0, 16), (2, 100.0, 16), (0, 4.2, 8), (1, 4.2, 32)])
def test_batch_equals_per_camera_traces(tile, tmax, kbuf):
    # a tmax of 4.2 from eyes at radius 4 cuts through the scene: exits beyond it follow the 16-event rounds
    print(f"case: 4 uneven cameras, tile {tile}, tmax {tmax}, K {kbuf}", flush=True)
    _check_batch(_scene_tensors(), CAMS, tile, tmax, kbuf, 0.0)


def test_batch_with_cameras_inside_primitives():
    # view 0 and view 2 sit inside primitives (tmin 0.05): each view's start-inside density and its adjoint come
    # from that view's own origin, as in a call of its camera alone
    print("case: 4 cameras, two inside primitives, tmin 0.05", flush=True)
    t = _scene_tensors(20_000, 16)
    means = t[0].cpu()
    cams = list(CAMS)
    for v, p in ((0, 11), (2, 4321)):
        # near the primitive's centre, not on it (its SH colour at its own centre has no direction)
        eye = tuple(float(x) + 0.004 for x in means[p])
        cams[v] = (eye, cams[v][1], cams[v][2])
    out = _check_batch(t, cams, 0, 100.0, 16, 0.05)
    # rays of the two inside views start with density, the others without
    start = 0
    for v, (_, W, H) in enumerate(cams):
        dens0 = float(out["initial_drgb"][start:start + W * H, 0].max())
        assert (dens0 > 0.0) == (v in (0, 2)), v
        start += W * H


def _check_batch(t, cams, tile, tmax, kbuf, tmin):
    step = B.BatchStep(t, cams, kbuf=kbuf, tile=tile, tmax=tmax, tmin=tmin)
    out = step.run()
    # per camera: its own single-view scene (SH at its eye) and its own trace
    # gradients are clamped per call (+-1e3, densities +-50): one call per camera clamps each camera's sum, the
    # batch clamps the sum over cameras, so elements that reach a clamp anywhere are left out of the comparison
    limit = {"dmeans": 1e3, "dscales": 1e3, "dquats": 1e3, "ddensities": 50.0, "dfeatures": 1e3}
    sums = {k: torch.zeros_like(out[k]) for k in limit}
    clamped = {k: out[k].abs() >= 0.999 * limit[k] for k in limit}
    start = 0
    for v, (eye, W, H) in enumerate(cams):
        nr = W * H
        sl = slice(start, start + nr)
        sc = T.Scene(*t, color_mode=1, sh_origin=eye, sh_degree=3, activation=1).build()
        o, d = step.rayo[sl].contiguous(), step.rayd[sl].contiguous()
        f = T.trace_forward(sc, o, d, tmin, tmax, 500, traversal="ray")
        if tmin == 0.0:
            assert torch.equal(f["color"], out["color"][sl]), v
        else:
            # the start-inside density is an FP64 sum over a candidate list built with atomics: its order, and so
            # the last bit of the sum, may differ between calls
            torch.testing.assert_close(f["color"], out["color"][sl], rtol=1e-6, atol=1e-7)
            torch.testing.assert_close(f["initial_drgb"], out["initial_drgb"][sl], rtol=1e-6, atol=1e-7)
        assert torch.equal(f["iters"], out["iters"][sl]), v
        b = T.trace_backward(sc, o, d, tmin, tmax, 500, f, step.grad[sl].contiguous())
        for k in sums:
            sums[k] += b[k]
            clamped[k] |= b[k].abs() >= 0.999 * limit[k]
        assert _rel(out["drayo"][sl], b["drayo"]) < 1e-6 and _rel(out["drayd"][sl], b["drayd"]) < 1e-6
        start += nr
    for k in sums:
        keep = ~clamped[k]
        err = _rel(out[k][keep], sums[k][keep])
        print(k, err, "clamped", int(clamped[k].sum()), "of", clamped[k].numel(), flush=True)
        assert err < 2e-5 and float(keep.float().mean()) > 0.99, (k, err)
    return out


def test_persistent_launch_equals_plain_launch():
    t = _scene_tensors(20_000, 13)
    a = B.BatchStep(t, CAMS, kbuf=16, tile=0).run()
    for tile in (1, 3):
        b = B.BatchStep(t, CAMS, kbuf=16, tile=tile).run()
        assert torch.equal(a["color"], b["color"]) and torch.equal(a["iters"], b["iters"])


def test_graph_replays_match_eager():
    print("case: hipGraph of build + forward + backward", flush=True)
    t = _scene_tensors(20_000, 14)
    eager = B.BatchStep(t, CAMS, kbuf=16, tile=2).run()
    ref = {k: v.clone() for k, v in eager.items() if torch.is_tensor(v)}
    step = B.BatchStep(t, CAMS, kbuf=16, tile=2).capture()
    for rep in range(3):
        out = step.replay()
        torch.cuda.synchronize()
        assert torch.equal(out["color"], ref["color"]) and torch.equal(out["iters"], ref["iters"]), rep
        for k in ("dmeans", "dscales", "dquats", "ddensities", "dfeatures"):
            assert _rel(out[k], ref[k]) < 1e-5, (rep, k)
        # eager work between replays
        B.BatchStep(t, CAMS[:2], kbuf=8, tile=0).run()
        torch.cuda.synchronize()


def test_autotuner_picks_a_measured_candidate(tmp_path):
    os.environ["EVER_AUTOTUNE_CACHE"] = str(tmp_path / "tune.json")
    t = _scene_tensors(10_000, 15)
    cands = [(8, 0), (16, 0), (16, 2)]
    key = B.workload_key(10_000, sum(w * h for _, w, h in CAMS), len(CAMS), D)
    best, times = B.autotune(lambda k, tl: B.BatchStep(t, CAMS, kbuf=k, tile=tl), cands, reps=2, key=key)
    assert best in cands and set(times) == set(cands) and times[best] == min(times.values())
    assert (tmp_path / "tune.json").exists()
    best2, times2 = B.autotune(lambda k, tl: None, cands, key=key)  # served from the cache
    assert best2 == best and times2 == times
