"""FP64 restatement of the reference caller's camera-ray generation — TEST INFRASTRUCTURE ONLY.

Only tests/ import this module, as the checker of ever_amd.cameras (ever_camera_rays). Follows
integration_utils/gaussian_renderer/ever.py:29-121 (get_ray_directions, get_rays, camera2rays, camera2rays_full) and
integration_utils/utils/camera_utils_zipnerf.py:816-906 (radial/tangential undistortion by Newton's method) and
:950-1030 (pixels_to_rays: pixel centre, inverse intrinsics, fisheye and panoramic mappings, rotation to world,
normalisation). Pinned in tests/test_cameras_oracle.py by reprojection, the equidistant-fisheye angle law, the
panoramic spherical angles and the distortion round trip, none of which re-evaluates these formulas.
"""
import numpy as np

PERSPECTIVE, FISHEYE, PANORAMIC = 0, 1, 2


def distort(x, y, k1=0.0, k2=0.0, k3=0.0, k4=0.0, p1=0.0, p2=0.0):
    """The forward lens model (camera_utils_zipnerf.py:836-839): undistorted -> distorted camera-plane point."""
    r = x * x + y * y
    d = 1.0 + r * (k1 + r * (k2 + r * (k3 + r * k4)))
    return d * x + 2 * p1 * x * y + p2 * (r + 2 * x * x), d * y + 2 * p2 * x * y + p1 * (r + 2 * y * y)


def undistort(xd, yd, k1=0.0, k2=0.0, k3=0.0, k4=0.0, p1=0.0, p2=0.0, eps=1e-9, iterations=10):
    """camera_utils_zipnerf.py:866-906: Newton's method on distort(x, y) = (xd, yd) from (xd, yd), a step skipped
    where |det J| <= eps."""
    x, y = np.array(xd, np.float64), np.array(yd, np.float64)
    for _ in range(iterations):
        r = x * x + y * y
        d = 1.0 + r * (k1 + r * (k2 + r * (k3 + r * k4)))
        ex = d * x + 2 * p1 * x * y + p2 * (r + 2 * x * x) - xd
        ey = d * y + 2 * p2 * x * y + p1 * (r + 2 * y * y) - yd
        dr = k1 + r * (2 * k2 + r * (3 * k3 + r * 4 * k4))
        dx, dy = 2 * x * dr, 2 * y * dr
        jxx = d + dx * x + 2 * p1 * y + 6 * p2 * x
        jxy = dy * x + 2 * p1 * x + 2 * p2 * y
        jyx = dx * y + 2 * p2 * y + 2 * p1 * x
        jyy = d + dy * y + 2 * p2 * x + 6 * p1 * y
        det = jxx * jyy - jxy * jyx
        ok = np.abs(det) > eps
        safe = np.where(ok, det, 1.0)
        x = x - np.where(ok, (ex * jyy - ey * jxy) / safe, 0.0)
        y = y - np.where(ok, (ey * jxx - ex * jyx) / safe, 0.0)
    return x, y


def camera_directions(u, v, model):
    """Camera-space direction of camera-plane points (camera_utils_zipnerf.py:981-1006)."""
    if model == FISHEYE:
        th = np.minimum(np.sqrt(u * u + v * v), np.pi)
        s = np.where(th > 0, np.sin(th) / np.where(th > 0, th, 1.0), 1.0)
        return np.stack([u * s, v * s, np.cos(th)], -1)
    if model == PANORAMIC:
        return np.stack([-np.sin(v) * np.sin(u), -np.cos(v), -np.sin(v) * np.cos(u)], -1)
    return np.stack([u, v, np.ones_like(u)], -1)


def camera_rays(width, height, fx, fy, c2w, cx=None, cy=None, model=PERSPECTIVE, distortion=None, jitter=None):
    """(origins, unit directions), each (width*height, 3) float64, pixels in row-major order."""
    c2w = np.asarray(c2w, np.float64)[:3, :4]
    cx = 0.5 * width if cx is None else cx
    cy = 0.5 * height if cy is None else cy
    y, x = np.meshgrid(np.arange(height, dtype=np.float64), np.arange(width, dtype=np.float64), indexing="ij")
    x, y = x.reshape(-1), y.reshape(-1)
    if jitter is None:
        jx = jy = 0.5
    else:
        j = np.asarray(jitter, np.float64).reshape(-1, 2)
        jx, jy = j[:, 0], j[:, 1]
    u = (x + jx - cx) / fx
    v = (y + jy - cy) / fy
    if distortion:
        u, v = undistort(u, v, **distortion)
    dc = camera_directions(u, v, model)
    dw = dc @ c2w[:, :3].T
    dw = dw / np.linalg.norm(dw, axis=1, keepdims=True)
    return np.broadcast_to(c2w[:, 3], dw.shape).copy(), dw
