"""The training renderer (ever_amd.render.splinerender, reference integration_utils/gaussian_renderer/ever.py:123-224)
with its GPU camera rays and fused SH prepass equals the reference caller's composition on the same path's
pieces: rays, eval_sh -> softplus(beta=10) -> RGB2SH, trace_rays with the means2D statistic; forward outputs,
densification statistics and gradients."""
import math
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from ever_amd import cameras as CAM
from ever_amd import scenes
from ever_amd.eval_sh import eval_sh
from ever_amd.render import splinerender
from ever_amd.splinetracers import fast_ellipsoid_splinetracer as fes

pytestmark = pytest.mark.gpu
D = torch.device("cuda:0")
C0 = 0.28209479177387814


class _Model:
    """The attributes of the reference's GaussianModel that the renderer reads."""

    def __init__(self, sc, deg):
        self._xyz = sc["means"].to(D).requires_grad_(True)
        self.features = sc["features"].to(D).requires_grad_(True)
        self.rot = sc["quats"].to(D).requires_grad_(True)
        self.scales = sc["scales"].to(D).requires_grad_(True)
        self.density = sc["densities"].to(D).requires_grad_(True)
        self.active_sh_degree = deg
        self.tmin = 0.0
        self.glo_network = None

    get_xyz = property(lambda self: self._xyz)
    get_features = property(lambda self: self.features)
    get_rotation = property(lambda self: self.rot)

    def get_scale_and_density_for_rendering(self, filter_scale, modifier):
        return self.scales * modifier, self.density

    def params(self):
        return [self._xyz, self.features, self.rot, self.scales, self.density]


def _view(eye, W, H, model="perspective", dist=None):
    Rcw, e = scenes.look_at(eye)
    w2c = torch.eye(4, dtype=torch.float64)
    w2c[:3, :3] = Rcw.double().T
    w2c[:3, 3] = -Rcw.double().T @ e.double()
    return SimpleNamespace(image_width=W, image_height=H, FoVx=1.0, FoVy=2 * math.atan(H / W * math.tan(0.5)),
                           world_view_transform=w2c.T.float().to(D), camera_center=e.float().to(D), model=model,
                           distortion_params=dist, glo_vector=None)


@pytest.mark.parametrize("model, dist", [("perspective", None), ("fisheye", dict(k1=0.03))])
def test_splinerender_equals_the_reference_composition(model, dist):
    print(f"case: splinerender {model}", flush=True)
    sc = scenes.uniform_scene(5000, 3, seed=31)
    view = _view((0.3, 0.4, -3.5), 96, 72, model, dist)
    pipe = SimpleNamespace(enable_GLO=False)
    g = torch.Generator().manual_seed(5)
    up = torch.randn((3, 72, 96), generator=g).to(D)

    pa = _Model(sc, 3)
    a = splinerender(view, pa, pipe, None)
    (a["render"] * up).sum().backward()

    pb = _Model(sc, 3)
    o, d = CAM.view_rays(view, device=D)
    W, H = 96, 72
    fx, fy = CAM.focal_from_fov(W, H, view.FoVx, view.FoVy)
    K = torch.tensor([[fx, 0, W / 2, 0], [0, fy, H / 2, 0], [0, 0, 1, 0]], device=D, dtype=torch.float32)
    full = torch.eye(4, device=D)
    full[:, :3] = view.world_view_transform @ K.T
    m2 = torch.zeros_like(pb._xyz[:, :2]).requires_grad_(True)
    rgb = torch.nn.functional.softplus(eval_sh(pb._xyz, pb.features, view.camera_center, 3), beta=10)
    f0 = ((rgb - 0.5) / C0).reshape(-1, 1, 3)
    out, ex = fes.trace_rays(pb._xyz, pb.scales, pb.rot, pb.density, f0, o, d, 0.0, 1e7, 100, m2,
                             full.reshape(1, 4, 4), max_iters=400, return_extras=True)
    render = out[:, :3].T.reshape(3, H, W)
    (render * up).sum().backward()

    assert float((a["render"] - render).detach().abs().max()) < 1e-5
    assert torch.equal(a["touch_count"], ex["touch_count"])
    assert torch.equal(a["iters"], ex["iters"].reshape(H, W))
    assert torch.equal(a["visibility_filter"], (ex["touch_count"] // 2) >= 4)
    assert a["opacity"].shape == (W * H, 1) and a["distortion_loss"].shape == (W * H, 1)
    assert float((a["opacity"][:, 0] - out[:, 3]).detach().abs().max()) < 1e-5
    for x, y, name in zip(pa.params() + [a["viewspace_points"]], pb.params() + [m2],
                          ("means", "features", "quats", "scales", "densities", "means2D")):
        rel = float((x.grad - y.grad).norm() / max(float(y.grad.norm()), 1e-12))
        print(name, rel, flush=True)
        assert rel < 1e-4, name


def test_fast_renderer_equals_the_sp_composition():
    # the reference's FastRenderer frame (fast_renderer.py:110-131) through the sp module with tensor-computed
    # colours, against the fused renderer, over two views of one model (the tree is built once)
    from ever_amd.render import FastRenderer
    from ever_amd.splinetracers.fast_ellipsoid_splinetracer import sp
    print("case: FastRenderer, two views", flush=True)
    sc = scenes.uniform_scene(5000, 3, seed=33)
    pc = _Model(sc, 3)
    views = [_view((0.3, 0.4, -3.5), 96, 72), _view((-2.0, 1.0, -2.5), 80, 60)]
    fr = FastRenderer(views[0], pc, False)
    with torch.no_grad():
        prims = sp.Primitives(D)
        half = torch.cat([pc._xyz, pc.scales, pc.rot], 1).half().contiguous()
        prims.add_primitives(pc._xyz.detach(), pc.scales.detach(), pc.rot.detach(), half, pc.density.detach(),
                             fr.get_color(views[0]))
        ctx = sp.OptixContext(D)
        gas = sp.GAS(ctx, D, prims, True, False, True)
        fwd = sp.Forward(ctx, D, prims, False)
        for view in views:
            img = fr.render(view, pc, None)
            prims.set_features(fr.get_color(view))
            fwd.update_model(prims)
            o, d = CAM.view_rays(view, device=D)
            ref = fwd.trace_rays(gas, o, d, 0.0, 1e7, 200, 1000)["color"]
            ref_img = ref[:, :3].T.reshape(3, view.image_height, view.image_width)
            err = float((img - ref_img).abs().max())
            print("max abs diff", err, flush=True)
            assert img.shape == (3, view.image_height, view.image_width) and err < 1e-5
