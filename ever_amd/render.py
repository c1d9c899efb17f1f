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

from . import cameras as CAM
from .splinetracers.fast_ellipsoid_splinetracer import trace_rays_sh

MAX_ITERS = 400  # ever.py:20


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
