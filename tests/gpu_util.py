"""Helpers shared by the GPU parity tests: run the HIP tracer and the FP64 oracle on the same inputs."""
import numpy as np
import torch

from ever_amd import tracer as T
from oracle import oracle

KEYS = ("means", "scales", "quats", "densities", "features")


def dev():
    return torch.device("cuda:0")


def gpu_forward(sc, o, d, tmin=0.0, tmax=100.0, max_iters=500, kbuf=None, traversal=None, **scene_kw):
    D = dev()
    t = [sc[k].to(D).contiguous() for k in KEYS]
    scene = T.Scene(*t, **scene_kw).build()
    out = T.trace_forward(scene, o.to(D).contiguous(), d.to(D).contiguous(), tmin, tmax, max_iters, kbuf=kbuf,
                          traversal=traversal)
    return scene, out


def oracle_forward(sc, o, d, tmin=0.0, tmax=100.0, max_iters=500, kbuf=16, use_bvh=False):
    return oracle.forward(*[sc[k].double().numpy() for k in KEYS], o.double().numpy(), d.double().numpy(), tmin,
                          tmax, max_iters=max_iters, kbuf=kbuf, use_bvh=use_bvh)


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-12)
