"""Camera rays generated on the GPU in one kernel (ever_camera_rays).

Replaces the reference caller's ray generation (integration_utils/gaussian_renderer/ever.py:29-121):
``get_ray_directions`` + ``get_rays`` for the pinhole camera (``camera2rays``, :104-121) and
``camera2rays_full`` (:77-102 -> integration_utils/utils/camera_utils_zipnerf.py:909-1040 ``pixels_to_rays``) for
fisheye and panoramic cameras with radial/tangential distortion. Rays come out in row-major pixel order
(r = y * width + x), origins = the camera centre, unit directions in world space, OpenCV axes (+z forward).

Deliberate difference: a fisheye pixel exactly on the optical axis gets the axis direction (the reference's
sin(theta)/theta is 0/0 there).
"""
import ctypes

import torch

from . import _capi as C

# the reference's ProjectionType values (scene.dataset_readers; camera_utils_zipnerf.py: 'perspective', 'fisheye',
# 'pano')
PROJECTIONS = {"perspective": C.EVER_CAMERA_PERSPECTIVE, "fisheye": C.EVER_CAMERA_FISHEYE,
               "pano": C.EVER_CAMERA_PANORAMIC, "panoramic": C.EVER_CAMERA_PANORAMIC}
DISTORTION_KEYS = ("k1", "k2", "k3", "k4", "p1", "p2")


def projection_code(model):
    """EVER_CAMERA_* code of a projection given as a name, an enum member (by value or name) or a code."""
    if model is None:
        return C.EVER_CAMERA_PERSPECTIVE
    if isinstance(model, int) and model in PROJECTIONS.values():
        return model
    for cand in (getattr(model, "value", None), getattr(model, "name", None), model):
        if isinstance(cand, str) and cand.lower() in PROJECTIONS:
            return PROJECTIONS[cand.lower()]
    raise ValueError(f"unknown camera projection {model!r}")


def camera_rays(width, height, fx, fy, c2w, cx=None, cy=None, model="perspective", distortion=None, jitter=None):
    """(origins, directions), each (width*height, 3) float32 on c2w's device.

    c2w: (3,4) or (4,4) camera-to-world matrix (a device tensor: nothing is read back to the host).
    cx, cy: principal point in pixels (default width/2, height/2, as ever.py:47 and :86-87).
    distortion: dict with any of k1, k2, k3, k4, p1, p2 (camera_utils_zipnerf.py:866-877), or None.
    jitter: (width*height, 2) or (height, width, 2) pixel offsets in [0, 1) replacing the pixel centre 0.5
    (get_ray_directions(random=True), ever.py:43-46)."""
    if not isinstance(c2w, torch.Tensor) or not c2w.is_cuda:
        raise RuntimeError("c2w must be a CUDA tensor")
    if c2w.dim() != 2 or c2w.size(0) < 3 or c2w.size(1) != 4:
        raise RuntimeError("c2w must be a (3,4) or (4,4) matrix")
    dev = c2w.device
    width, height = int(width), int(height)
    n = width * height
    M = c2w[:3, :4].to(torch.float32).contiguous()
    cam = C.Camera()
    cam.width, cam.height = width, height
    cam.model = projection_code(model)
    cam.fx, cam.fy = float(fx), float(fy)
    cam.cx = 0.5 * width if cx is None else float(cx)
    cam.cy = 0.5 * height if cy is None else float(cy)
    if distortion:
        unknown = set(distortion) - set(DISTORTION_KEYS)
        if unknown:
            raise ValueError(f"unknown distortion parameters {sorted(unknown)}")
        cam.distorted = 1
        for i, k in enumerate(DISTORTION_KEYS):
            cam.dist[i] = float(distortion.get(k, 0.0))
    cam.c2w = M.data_ptr()
    if jitter is not None:
        if not (isinstance(jitter, torch.Tensor) and jitter.device == dev and jitter.dtype == torch.float32
                and jitter.is_contiguous() and jitter.numel() == 2 * n and jitter.size(-1) == 2):
            raise RuntimeError("jitter must be a contiguous float32 (H*W,2) tensor on the camera's device")
    origins = torch.empty((n, 3), dtype=torch.float32, device=dev)
    directions = torch.empty((n, 3), dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        C.check(C.lib().ever_camera_rays(ctypes.byref(cam), C.ptr(jitter), C.ptr(origins), C.ptr(directions),
                                         C.stream_ptr()), "ever_camera_rays")
    return origins, directions


def focal_from_fov(width, height, fovx, fovy):
    """fx, fy = 0.5 W / tan(FoVx / 2), 0.5 H / tan(FoVy / 2) (ever.py:111-112)."""
    import math
    return 0.5 * width / math.tan(0.5 * fovx), 0.5 * height / math.tan(0.5 * fovy)


def view_c2w(view, device=None):
    """Camera-to-world matrix of a view with a (4,4) world_view_transform (the transpose of world-to-camera):
    inverse of its transpose (ever.py:117), computed on the device."""
    wvt = view.world_view_transform
    if device is not None:
        wvt = wvt.to(device)
    return torch.linalg.inv(wvt.T.float())


def camera2rays(view, random=False, generator=None, device=None):
    """Pinhole rays of a view (ever.py:104-121); random=True jitters each ray inside its pixel."""
    w, h = view.image_width, view.image_height
    fx, fy = focal_from_fov(w, h, view.FoVx, view.FoVy)
    c2w = view_c2w(view, device or torch.device("cuda"))
    jitter = None
    if random:
        jitter = torch.rand((h * w, 2), generator=generator, device=c2w.device)
    return camera_rays(w, h, fx, fy, c2w, jitter=jitter)


def camera2rays_full(view, device=None, **kwargs):
    """Rays of a view of any projection with its distortion_params (ever.py:77-102); pixel centres."""
    w, h = view.image_width, view.image_height
    fx, fy = focal_from_fov(w, h, view.FoVx, view.FoVy)
    c2w = view_c2w(view, device or torch.device("cuda"))
    return camera_rays(w, h, fx, fy, c2w, model=getattr(view, "model", None),
                       distortion=getattr(view, "distortion_params", None))


def view_rays(view, random=False, device=None):
    """The reference renderer's choice (ever.py:135-138): pinhole rays for perspective views, else the full
    camera model."""
    if projection_code(getattr(view, "model", None)) == C.EVER_CAMERA_PERSPECTIVE:
        return camera2rays(view, random=random, device=device)
    return camera2rays_full(view, device=device)
