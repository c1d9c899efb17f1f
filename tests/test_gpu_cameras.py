"""Camera rays from the GPU kernel (ever_camera_rays) against the FP64 camera oracle: pinhole, fisheye and
panoramic projections, with and without lens distortion and pixel jitter, odd and empty image sizes, the
view-object helpers (ever.py:77-121), and a 1920x1080 camera through size-independent properties (unit
directions, reprojection to the pixel centres)."""
import math
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from ever_amd import cameras as CAM
from oracle import cameras as K

pytestmark = pytest.mark.gpu
D = torch.device("cuda:0")
TOL = 5e-6  # float32 kernel vs FP64 oracle, unit directions


def _c2w(seed):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=4)
    w, x, y, z = q / np.linalg.norm(q)
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    return np.concatenate([R, rng.normal(size=(3, 1))], 1).astype(np.float32)


CASES = [("perspective", None), ("perspective", dict(k1=-0.2, k2=0.05, p1=1e-3, p2=-2e-3)),
         ("fisheye", None), ("fisheye", dict(k1=0.05, k2=-0.01, k3=1e-3, k4=-1e-4)), ("pano", None)]


@pytest.mark.parametrize("model, dist", CASES)
@pytest.mark.parametrize("jit", [False, True])
def test_rays_match_oracle(model, dist, jit):
    print(f"case: camera {model} distortion {dist is not None} jitter {jit}", flush=True)
    W, H = 97, 61
    fx, fy = (W / (2 * math.pi), H / math.pi) if model == "pano" else (70.0, 66.0)
    cy = 0.0 if model == "pano" else None
    M = _c2w({"perspective": 7, "fisheye": 8, "pano": 9}[model])
    jitter = torch.rand((W * H, 2), generator=torch.Generator().manual_seed(3)) if jit else None
    o, d = CAM.camera_rays(W, H, fx, fy, torch.from_numpy(M).to(D), cy=cy, model=model, distortion=dist,
                           jitter=None if jitter is None else jitter.to(D))
    ro, rd = K.camera_rays(W, H, fx, fy, M.astype(np.float64), cy=cy, model=CAM.projection_code(model),
                           distortion=dist, jitter=None if jitter is None else jitter.double().numpy())
    assert o.shape == (W * H, 3) and d.shape == (W * H, 3)
    assert np.abs(o.cpu().numpy() - ro).max() == 0.0
    err = np.abs(d.cpu().double().numpy() - rd).max()
    print("max direction error", err, flush=True)
    assert err < TOL


def test_odd_and_empty_sizes():
    M = torch.from_numpy(_c2w(11)).to(D)
    for W, H in ((1, 1), (7, 5), (0, 5), (5, 0)):
        o, d = CAM.camera_rays(W, H, 3.0, 3.0, M)
        assert o.shape == (W * H, 3)
        if W * H:
            _, rd = K.camera_rays(W, H, 3.0, 3.0, M.cpu().double().numpy())
            assert np.abs(d.cpu().double().numpy() - rd).max() < TOL


def test_view_helpers_match_the_reference_conventions():
    # a 3DGS-style view: world_view_transform = (world-to-camera)^T, FoV in radians
    M = _c2w(12).astype(np.float64)
    w2c = np.eye(4)
    w2c[:3, :3] = M[:, :3].T
    w2c[:3, 3] = -M[:, :3].T @ M[:, 3]
    view = SimpleNamespace(image_width=80, image_height=60, FoVx=1.1, FoVy=0.9, model="perspective",
                           world_view_transform=torch.tensor(w2c.T, dtype=torch.float32, device=D),
                           distortion_params=None)
    fx, fy = 0.5 * 80 / math.tan(0.55), 0.5 * 60 / math.tan(0.45)
    _, rd = K.camera_rays(80, 60, fx, fy, M)
    for fn in (CAM.camera2rays, CAM.camera2rays_full, CAM.view_rays):
        o, d = fn(view)
        assert np.abs(o.cpu().numpy() - M[:, 3]).max() < 1e-5
        assert np.abs(d.cpu().double().numpy() - rd).max() < 1e-5
    view.model = "fisheye"
    view.distortion_params = dict(k1=0.02)
    _, rd = K.camera_rays(80, 60, fx, fy, M, model=K.FISHEYE, distortion=dict(k1=0.02))
    _, d = CAM.view_rays(view)
    assert np.abs(d.cpu().double().numpy() - rd).max() < 1e-5
    # jittered pinhole rays stay inside their pixels: reproject
    o, d = CAM.camera2rays(view, random=True, generator=torch.Generator(device=D).manual_seed(1))
    c = d.cpu().double().numpy() @ M[:, :3]
    px, py = fx * c[:, 0] / c[:, 2] + 40, fy * c[:, 1] / c[:, 2] + 30
    y, x = np.divmod(np.arange(80 * 60), 80)
    assert ((px - x) > -1e-3).all() and ((px - x) < 1 + 1e-3).all()
    assert ((py - y) > -1e-3).all() and ((py - y) < 1 + 1e-3).all()


def test_full_hd_properties():
    print("case: camera 1920x1080 properties", flush=True)
    W, H, f = 1920, 1080, 1662.0
    M = _c2w(13)
    o, d = CAM.camera_rays(W, H, f, f, torch.from_numpy(M).to(D))
    d64 = d.double()
    assert float((d64.norm(dim=1) - 1).abs().max()) < 1e-6
    Rt = torch.from_numpy(M[:, :3]).double().to(D)
    c = d64 @ Rt
    px, py = f * c[:, 0] / c[:, 2] + W / 2, f * c[:, 1] / c[:, 2] + H / 2
    r = torch.arange(W * H, device=D)
    assert float((px - (r % W + 0.5)).abs().max()) < 2e-3 and float((py - (r // W + 0.5)).abs().max()) < 2e-3
