"""Pins the camera-ray oracle (oracle/cameras.py) on the CPU by properties that do not re-evaluate its formulas:
pinhole reprojection to the pixel, the principal ray along the camera's +z axis, the equidistant-fisheye angle
law (angle to the axis = |camera-plane point|, azimuth kept), the panoramic spherical angles, and the lens model's
round trip distort(undistort(p)) = p. Scenario source: integration_utils/gaussian_renderer/ever.py:29-121 and
integration_utils/utils/camera_utils_zipnerf.py:816-1030."""
import numpy as np
import pytest

from oracle import cameras as K


def _rot(seed):
    q = np.random.default_rng(seed).normal(size=4)
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _c2w(seed):
    R = _rot(seed)
    t = np.random.default_rng(seed + 100).normal(size=3)
    return np.concatenate([R, t[:, None]], 1), R, t


@pytest.mark.parametrize("jit", [False, True])
def test_pinhole_rays_reproject_to_their_pixels(jit):
    W, H, fx, fy, cx, cy = 37, 23, 41.0, 39.5, 17.2, 12.9
    M, R, t = _c2w(1)
    jitter = np.random.default_rng(2).random((W * H, 2)) if jit else None
    o, d = K.camera_rays(W, H, fx, fy, M, cx=cx, cy=cy, jitter=jitter)
    assert np.allclose(o, t) and np.allclose(np.linalg.norm(d, axis=1), 1.0)
    c = d @ R  # camera space (R^T d)
    assert (c[:, 2] > 0).all()
    px = fx * c[:, 0] / c[:, 2] + cx
    py = fy * c[:, 1] / c[:, 2] + cy
    y, x = np.divmod(np.arange(W * H), W)
    j = jitter if jit else np.full((W * H, 2), 0.5)
    assert np.abs(px - (x + j[:, 0])).max() < 1e-9 and np.abs(py - (y + j[:, 1])).max() < 1e-9


def test_principal_ray_is_the_optical_axis():
    M, R, _ = _c2w(3)
    for model in (K.PERSPECTIVE, K.FISHEYE):
        # a 1x1 image centred at its pixel centre: the one ray is R e_z
        _, d = K.camera_rays(1, 1, 10.0, 10.0, M, cx=0.5, cy=0.5, model=model)
        assert np.allclose(d[0], R[:, 2], atol=1e-12)


def test_fisheye_angle_law():
    W, H, f = 64, 48, 12.0  # camera-plane radius up to ~3.3 rad: past pi, where the angle is clipped
    M, R, _ = _c2w(4)
    _, d = K.camera_rays(W, H, f, f, M, model=K.FISHEYE)
    c = d @ R
    y, x = np.divmod(np.arange(W * H), W)
    u, v = (x + 0.5 - W / 2) / f, (y + 0.5 - H / 2) / f
    rho = np.hypot(u, v)
    ang = np.arccos(np.clip(c[:, 2], -1, 1))
    ok = rho < np.pi - 1e-6
    assert np.abs(ang[ok] - rho[ok]).max() < 1e-9
    assert np.allclose(c[~ok], [0, 0, -1], atol=1e-9)  # clipped at pi: straight back
    az = np.arctan2(c[ok, 1], c[ok, 0])
    assert np.abs(np.angle(np.exp(1j * (az - np.arctan2(v[ok], u[ok]))))).max() < 1e-9


def test_panoramic_spherical_angles():
    W, H = 40, 20
    M, R, _ = _c2w(5)
    fx, fy = W / (2 * np.pi), H / np.pi  # u in [-pi, pi), v in [-pi/2, pi/2)
    _, d = K.camera_rays(W, H, fx, fy, M, cx=W / 2, cy=0.0, model=K.PANORAMIC)
    c = d @ R
    y, x = np.divmod(np.arange(W * H), W)
    u, v = (x + 0.5 - W / 2) / fx, (y + 0.5) / fy  # v = polar angle from -y, in (0, pi)
    assert np.abs(np.arccos(np.clip(-c[:, 1], -1, 1)) - v).max() < 1e-9
    az = np.arctan2(-c[:, 0], -c[:, 2])
    assert np.abs(np.angle(np.exp(1j * (az - u)))).max() < 1e-9


@pytest.mark.parametrize("coef", [dict(k1=-0.25, k2=0.08), dict(k1=0.1, k2=-0.02, k3=0.003, k4=-1e-4, p1=2e-3,
                                                                p2=-1.5e-3), dict(p1=0.01, p2=0.01)])
def test_undistortion_inverts_the_lens_model(coef):
    rng = np.random.default_rng(6)
    xd, yd = rng.uniform(-0.8, 0.8, 2000), rng.uniform(-0.6, 0.6, 2000)
    x, y = K.undistort(xd, yd, **coef)
    bx, by = K.distort(x, y, **coef)
    assert np.abs(bx - xd).max() < 1e-10 and np.abs(by - yd).max() < 1e-10
    # no coefficients: identity
    x0, y0 = K.undistort(xd, yd)
    assert np.array_equal(x0, xd) and np.array_equal(y0, yd)


def test_distorted_rays_reproject_through_the_lens_model():
    W, H, f = 30, 20, 25.0
    coef = dict(k1=-0.2, k2=0.05, p1=1e-3, p2=-2e-3)
    M, R, _ = _c2w(7)
    _, d = K.camera_rays(W, H, f, f, M, distortion=coef)
    c = d @ R
    bx, by = K.distort(c[:, 0] / c[:, 2], c[:, 1] / c[:, 2], **coef)
    y, x = np.divmod(np.arange(W * H), W)
    assert np.abs(f * bx + W / 2 - (x + 0.5)).max() < 1e-8 and np.abs(f * by + H / 2 - (y + 0.5)).max() < 1e-8
