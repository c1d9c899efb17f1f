"""The reference training loop's renderer (integration_utils/gaussian_renderer/ever.py:123-224 ``splinerender``) on
the GPU path: camera rays from one kernel (ever_amd.cameras) and the SH colour prepass with its softplus(beta=10)
fused into the scene build (trace_rays_sh), instead of the caller-side eval_sh -> softplus -> RGB2SH -> trace_rays
round trip. Same arguments, same returned dict; the model and view objects are used through the attributes the
reference reads (pc.get_xyz, get_features, get_rotation, active_sh_degree, tmin, glo_network, _xyz,
get_scale_and_density_for_rendering; view.image_width/height, FoVx/FoVy, world_view_transform, camera_center,
model, distortion_params, glo_vector).
"""
import math

import torch

from . import _capi as C
from . import cameras as CAM
from . import tracer as T
from .splinetracers.fast_ellipsoid_splinetracer import trace_rays_sh

MAX_ITERS = 400  # ever.py:20
FAST_MAX_ITERS = 200  # fast_renderer.py:38


def splinerender(view, pc, pipe, bg_color, scaling_modifier=1.0, override_color=None, random=False, tmin=None,
                 tmax=1e7):
    """ever.py:123-224. bg_color and override_color are accepted and unused, as in the reference."""
    device = pc.get_xyz.device
    rays_o, rays_d = CAM.view_rays(view, random=random, device=device)

    means2D = torch.zeros_like(pc.get_xyz[..., :2])
    means2D.requires_grad = True

    w, h = view.image_width, view.image_height
    fx, fy = CAM.focal_from_fov(w, h, view.FoVx, view.FoVy)
    # projection used by the means2D statistic (ever.py:147-163)
    K = torch.tensor([[fx, 0, w / 2, 0], [0, fy, h / 2, 0], [0, 0, 1, 0]], device=device, dtype=torch.float32)
    wct = view.world_view_transform.to(device).float()
    full_wct = torch.eye(4, device=device)
    full_wct[:, :3] = wct @ K.T

    shs = pc.get_features
    if getattr(pipe, "enable_GLO", False):
        glo_vector = view.glo_vector if view.glo_vector is not None else torch.zeros((1, 64), device=device)
        shs = pc.glo_network(glo_vector.reshape(1, -1), shs.reshape(shs.shape[0], -1)).reshape(shs.shape)

    cam_pos = view.camera_center
    per_point_2d_filter_scale = torch.zeros(pc._xyz.shape[0], device=pc._xyz.device)
    scales, density = pc.get_scale_and_density_for_rendering(per_point_2d_filter_scale, scaling_modifier)
    tmin = pc.tmin if tmin is None else tmin
    out, extras = trace_rays_sh(pc.get_xyz, scales, pc.get_rotation, density, shs, rays_o, rays_d,
                                sh_origin=cam_pos, sh_degree=pc.active_sh_degree, tmin=tmin, tmax=tmax,
                                max_prim_size=100, dL_dmeans2D=means2D, wcts=full_wct.reshape(1, 4, 4),
                                max_iters=MAX_ITERS, return_extras=True, activation="softplus")

    num_pixels = extras["touch_count"] // 2
    side_length = num_pixels.float().sqrt()
    radii = side_length / 2 * math.sqrt(2) * 2.5 * 5
    return {
        "render": out[:, :3].T.reshape(3, h, w),
        "viewspace_points": means2D,
        "visibility_filter": num_pixels >= 4,
        "touch_count": extras["touch_count"],
        "radii": radii,
        "iters": extras["iters"].reshape(h, w),
        "opacity": out[:, 3].reshape(-1, 1),
        "distortion_loss": out[:, 4].reshape(-1, 1),
    }


def _glo_features(pc, view, enable_glo):
    """The model's SH features, passed through its GLO network for the view when enabled (ever.py:165-173,
    fast_renderer.py:85-94)."""
    shs = pc.get_features
    if enable_glo:
        glo_vector = view.glo_vector if view.glo_vector is not None else torch.zeros((1, 64), device=shs.device)
        shs = pc.glo_network(glo_vector.reshape(1, -1), shs.reshape(shs.shape[0], -1)).reshape(shs.shape)
    return shs.contiguous()


def _camera_centre(view):
    c = view.camera_center
    return [float(v) for v in (c.detach().reshape(3).cpu() if isinstance(c, torch.Tensor) else c)]


class FastRenderer:
    """The reference's inference renderer (integration_utils/gaussian_renderer/fast_renderer.py:40-131) on the
    GPU path: the tree is built once from the model; each frame rewrites only the primitives' colours, with the
    SH evaluation at the camera, softplus(beta=10) and the degree-0 encoding fused into one pass over the
    primitives (the reference evaluates them with tensor operations and sets them as new features), generates the
    camera's rays in one kernel and traces without keeping an event list.

    Difference: for a non-perspective view the reference's get_rays rotates the world-space directions of
    camera2rays_full a second time (fast_renderer.py:68-71 with :101-108); here every view's rays come from its
    camera model once."""

    def __init__(self, view, pc, enable_GLO):
        self.device = pc.get_xyz.device
        self.enable_GLO = enable_GLO
        self.pc = pc
        self.mean = pc.get_xyz.contiguous()
        self.quat = pc.get_rotation.contiguous()
        self.per_point_2d_filter_scale = 1
        scales, density = pc.get_scale_and_density_for_rendering(self.per_point_2d_filter_scale, 1.0)
        self.scales, self.density = scales.contiguous(), density.contiguous()
        shs = _glo_features(pc, view, enable_GLO)
        self.scene = T.Scene(self.mean, self.scales, self.quat, self.density, shs, color_mode=C.EVER_COLOR_SH,
                             sh_origin=_camera_centre(view), sh_degree=pc.active_sh_degree, activation=1).build()
        self.view = view

    def set_camera(self, view):
        """fast_renderer.py:68-82: the next frames' camera (its rays are generated per frame)."""
        self.view = view

    def get_color(self, view):
        """fast_renderer.py:84-99: the degree-0 colour features RGB2SH(softplus_10(SH at the camera)), (N,1,3),
        as a tensor (the frame path computes them fused into the scene instead)."""
        from .eval_sh import eval_sh
        cam = torch.tensor(_camera_centre(view), dtype=torch.float32, device=self.device)
        rgb = torch.nn.functional.softplus(eval_sh(self.mean, _glo_features(self.pc, view, self.enable_GLO), cam,
                                                   self.pc.active_sh_degree), beta=10)
        return ((rgb - 0.5) / 0.28209479177387814).reshape(-1, 1, 3).contiguous()

    def get_rays(self, view):
        """fast_renderer.py:101-108: (rays_o, rays_d) of the view, (H*W, 3) each."""
        return CAM.view_rays(view, device=self.device)

    def trace_rays(self, rayo, rayd, view, tmin, tmax):
        """fast_renderer.py:110-117: recolour the scene for the view, then trace. Returns the forward's dict:
        color (R,4), iters, states, initial_drgb."""
        with torch.no_grad():
            self.scene.update_features(_glo_features(self.pc, view, self.enable_GLO), sh_origin=_camera_centre(view),
                                       sh_degree=self.pc.active_sh_degree)
            out = T.trace_forward(self.scene, rayo.contiguous(), rayd.contiguous(), tmin, tmax, FAST_MAX_ITERS,
                                  store_events=False, count_touches=False)
        return dict(color=out["color"][:, :4].contiguous(), iters=out["iters"], states=out["states"],
                    initial_drgb=out["initial_drgb"])

    def render(self, view, pc, bg_color, tmin=None, scaling_modifier=1.0):
        """fast_renderer.py:119-131: the (3, H, W) image of the view."""
        rays_o, rays_d = self.get_rays(view)
        out = self.trace_rays(rays_o, rays_d, view, self.pc.tmin if tmin is None else tmin, 1e7)
        return out["color"][:, :3].T.reshape(3, view.image_height, view.image_width)
