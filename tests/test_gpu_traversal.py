"""The two forward tree walks (one lane per ray, or the 64 rays of a wave together) gather the same k-buffer
events, so every forward output must agree bit for bit; checked on coherent camera rays, incoherent random rays,
ray counts that do not fill the last wave, start-inside rays, the max_iters / tmax limits, degenerate scenes and
the full 1,000,000-ellipsoid 1920x1080 workload."""
import numpy as np
import pytest
import torch

from ever_amd import scenes
from gpu_util import gpu_forward

pytestmark = pytest.mark.gpu

OUTPUTS = ("color", "states", "iters", "touch_count", "initial_drgb", "ray_state64")


def _same(a, b, max_iters):
    for k in OUTPUTS:
        assert torch.equal(a[k], b[k]), k
    nr = a["iters"].shape[0]
    ta = a["tri_collection"].view(max_iters, nr)
    tb = b["tri_collection"].view(max_iters, nr)
    mask = torch.arange(max_iters, device=ta.device)[:, None] < a["iters"][None, :]
    assert torch.equal(ta[mask], tb[mask])


def _both(sc, o, d, tmin=0.0, tmax=100.0, max_iters=500, kbuf=None, **kw):
    _, a = gpu_forward(sc, o, d, tmin, tmax, max_iters, kbuf=kbuf, traversal="ray", **kw)
    _, b = gpu_forward(sc, o, d, tmin, tmax, max_iters, kbuf=kbuf, traversal="wave", **kw)
    _same(a, b, max_iters)
    return a


@pytest.mark.parametrize("kbuf", [8, 16, 32])
def test_camera_rays(kbuf):
    print(f"case: config1 camera rays kbuf {kbuf}", flush=True)
    sc, o, d = scenes.config1()
    a = _both(sc, o, d, kbuf=kbuf)
    assert int(a["iters"].max()) > 0


def test_incoherent_rays_and_partial_wave():
    print("case: random rays, 1000 rays", flush=True)
    sc = scenes.uniform_scene(2000, 0, seed=11)
    g = torch.Generator().manual_seed(11)
    o = torch.rand((1000, 3), generator=g) * 4 - 2
    d = torch.randn((1000, 3), generator=g)
    a = _both(sc, o, d)
    assert int(a["iters"].max()) > 0


def test_lone_ray_in_last_wave():
    # 64 + 1 rays: the second wave holds one ray whose walk pushes far deeper than one stack entry per live lane
    # (the wave's stack keeps entry i in lane i, so lanes without a ray must stay in the wave)
    print("case: 65 camera rays, deep tree", flush=True)
    sc = scenes.clustered_scene(20_000, 0, seed=23)
    o, d = scenes.pinhole_rays(13, 5, (0.0, 0.0, -4.0))
    a = _both(sc, o, d)
    assert int(a["iters"].max()) > 0


def test_start_inside_limits_and_sh():
    print("case: start inside, limits, SH", flush=True)
    sc = scenes.uniform_scene(300, 0, seed=4)
    rng = np.random.default_rng(4)
    o = torch.tensor(np.tile(sc["means"][0].numpy(), (200, 1)) + rng.normal(0, 1e-3, (200, 3)), dtype=torch.float32)
    d = torch.tensor(rng.normal(size=(200, 3)), dtype=torch.float32)
    _both(sc, o, d, tmin=0.01)
    sc, o, d = scenes.config1()
    for mi, tmax in ((3, 100.0), (500, 3.0), (0, 100.0)):
        _both(sc, o, d, tmax=tmax, max_iters=mi)
    sc = scenes.clustered_scene(20_000, 3, seed=6)
    o, d = scenes.pinhole_rays(128, 96, (0.0, 0.0, -4.0))
    _both(sc, o, d, color_mode=1, sh_origin=(0.0, 0.0, -4.0), sh_degree=3, activation=1)


def test_degenerate_scenes():
    print("case: one primitive / none", flush=True)
    sc = scenes.uniform_scene(1, 0, seed=2)
    sc["means"][:] = torch.tensor([0.0, 0.0, 0.0])
    sc["scales"][:] = torch.tensor([0.5, 0.4, 0.3])
    o, d = scenes.pinhole_rays(8, 8, (0.0, 0.0, -3.0))
    a = _both(sc, o, d)
    assert int(a["iters"].max()) == 2
    sc0 = {k: v[:0] for k, v in scenes.uniform_scene(4, 0, seed=2).items()}
    a = _both(sc0, o, d)
    assert int(a["iters"].max()) == 0


def test_full_workload_1m_1080p():
    # the sorted full-size batch: AUTO takes the per-wave walk here
    print("case: 1M clustered, 1920x1080", flush=True)
    sc, o, d = scenes.config3()
    eye = (0.0, 0.0, -4.0)
    kw = dict(color_mode=1, sh_origin=eye, sh_degree=3, activation=1)
    _, a = gpu_forward(sc, o, d, traversal="ray", **kw)
    _, b = gpu_forward(sc, o, d, traversal="auto", **kw)
    _same(a, b, 500)
