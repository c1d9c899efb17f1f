"""Seeded synthetic scenes and pinhole cameras for the benchmark workloads and the parity tests.

Tensors are drawn on the CPU with torch.Generator (fp32) so the GPU and the CPU oracle see identical bits.
Camera convention as the reference's caller (integration_utils/gaussian_renderer/ever.py:29-75): pixel centres
at +0.5, OpenCV axes (x right, y down, z forward), directions normalized.
"""
import math

import torch

SH_C0 = 0.28209479177387814


def _logu(g, shape, lo, hi):
    return torch.exp(torch.empty(shape).uniform_(math.log(lo), math.log(hi), generator=g))


def _unit_quats(g, n):
    q = torch.randn((n, 4), generator=g)
    return q / q.norm(dim=1, keepdim=True)


def uniform_scene(n, sh_degree=0, seed=0, dens_hi=50.0):
    """Means U[-1,1]^3, scales logU[0.02,0.2], quats normalize(N(0,I4)), densities logU[0.1,dens_hi];
    degree 0: features N(0,0.5); degree 3: DC N(0,1), rest N(0,0.3)."""
    g = torch.Generator().manual_seed(seed)
    means = torch.empty((n, 3)).uniform_(-1.0, 1.0, generator=g)
    scales = _logu(g, (n, 3), 0.02, 0.2)
    quats = _unit_quats(g, n)
    dens = _logu(g, (n,), 0.1, dens_hi)
    d = (sh_degree + 1) ** 2
    if sh_degree == 0:
        feats = torch.randn((n, 1, 3), generator=g) * 0.5
    else:
        feats = torch.randn((n, d, 3), generator=g) * 0.3
        feats[:, 0, :] = torch.randn((n, 3), generator=g)
    return dict(means=means, scales=scales, quats=quats, densities=dens, features=feats)


def clustered_scene(n, sh_degree=3, seed=0, components=64, std=0.15, box=2.0, max_aspect=100.0):
    """64-component Gaussian mixture in [-2,2]^3, anisotropic scales base logU[0.005,0.05] times per-axis
    aspect factors (max/min <= max_aspect), densities logU[0.1,200], SH degree 3."""
    g = torch.Generator().manual_seed(seed)
    centres = torch.empty((components, 3)).uniform_(-box, box, generator=g)
    comp = torch.randint(0, components, (n,), generator=g)
    means = (centres[comp] + torch.randn((n, 3), generator=g) * std).clamp(-box, box)
    base = _logu(g, (n, 1), 0.005, 0.05)
    aspect = torch.exp(torch.empty((n, 3)).uniform_(0.0, math.log(max_aspect), generator=g))
    aspect = aspect / aspect.min(dim=1, keepdim=True).values
    aspect = aspect.clamp(max=max_aspect)
    scales = base * aspect / aspect.max(dim=1, keepdim=True).values.sqrt()
    quats = _unit_quats(g, n)
    dens = _logu(g, (n,), 0.1, 200.0)
    d = (sh_degree + 1) ** 2
    feats = torch.randn((n, d, 3), generator=g) * 0.3
    feats[:, 0, :] = torch.randn((n, 3), generator=g)
    return dict(means=means, scales=scales, quats=quats, densities=dens, features=feats)


def look_at(eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)):
    """Camera-to-world rotation (columns = camera x, y, z axes in world) for an OpenCV camera at eye."""
    eye = torch.tensor(eye, dtype=torch.float64)
    f = torch.tensor(target, dtype=torch.float64) - eye
    f = f / f.norm()
    upv = torch.tensor(up, dtype=torch.float64)
    if abs(float(torch.dot(f, upv))) > 0.999:
        upv = torch.tensor((0.0, 0.0, 1.0), dtype=torch.float64)
    x = torch.linalg.cross(f, upv)
    x = x / x.norm()
    y = torch.linalg.cross(f, x)
    return torch.stack([x, y, f], dim=1), eye


def pinhole_rays(width, height, eye, fov_deg=60.0, target=(0.0, 0.0, 0.0)):
    """(R,3) origins and unit directions, row-major pixels, focal = 0.5 W / tan(fov/2) on both axes."""
    Rcw, e = look_at(eye, target)
    f = 0.5 * width / math.tan(0.5 * math.radians(fov_deg))
    j, i = torch.meshgrid(torch.arange(height, dtype=torch.float64), torch.arange(width, dtype=torch.float64),
                          indexing="ij")
    dirs = torch.stack([(i + 0.5 - width / 2) / f, (j + 0.5 - height / 2) / f, torch.ones_like(i)], -1)
    dirs = dirs / dirs.norm(dim=-1, keepdim=True)
    d = (dirs.reshape(-1, 3) @ Rcw.T).float().contiguous()
    o = e.float().reshape(1, 3).expand(d.shape[0], 3).contiguous()
    return o, d


def world_to_camera(eye, target=(0.0, 0.0, 0.0)):
    """4x4 world-to-view matrix in the row-vector convention p_view = (p, 1) @ W (backwards_kernel.slang:244-250)."""
    Rcw, e = look_at(eye, target)
    W = torch.eye(4, dtype=torch.float64)
    W[:3, :3] = Rcw
    W[3, :3] = -(e.reshape(1, 3) @ Rcw).reshape(3)
    return W.float()


def config1(seed=0):
    """1,000 ellipsoids, SH0, one 64x64 camera at radius 3, tmin 0, tmax 100."""
    sc = uniform_scene(1000, 0, seed)
    o, d = pinhole_rays(64, 64, (0.0, 0.0, -3.0))
    return sc, o, d


def config2(seed=0, width=1280, height=720):
    """300,000 ellipsoids, SH3, one 1280x720 camera at radius 3."""
    sc = uniform_scene(300_000, 3, seed)
    o, d = pinhole_rays(width, height, (0.0, 0.0, -3.0))
    return sc, o, d


def config3(seed=0, width=1920, height=1080, n=1_000_000):
    """1,000,000 clustered ellipsoids, SH3, one 1920x1080 camera at radius 4."""
    sc = clustered_scene(n, 3, seed)
    o, d = pinhole_rays(width, height, (0.0, 0.0, -4.0))
    return sc, o, d


def upstream_grad(nr, seed=1):
    g = torch.Generator().manual_seed(seed)
    return torch.randn((nr, 5), generator=g)
