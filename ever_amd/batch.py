"""Several cameras of uneven resolution traced as one batch (not in the reference, where each camera is its own
trace_rays call): one ray batch with a per-ray view index, per-view SH colours in one scene build, the per-wave
walk in its persistent form, the whole build + forward + backward captured in one HIP graph, and a small
autotuner that picks the k-buffer size and the persistent tile size for a workload by timing them.
"""
import json
import math
import os
import time

import torch

from . import tracer as T


def fibonacci_eyes(n, radius):
    """n points on the sphere of the given radius (golden-angle spiral)."""
    out = []
    for i in range(n):
        y = 1.0 - 2.0 * (i + 0.5) / n
        r = math.sqrt(max(0.0, 1.0 - y * y))
        phi = i * math.pi * (3.0 - math.sqrt(5.0))
        out.append((radius * r * math.cos(phi), radius * y, radius * r * math.sin(phi)))
    return out


def uneven_cameras(n=16, radius=4.0, seed=3, w=(256, 1920), h=(256, 1080)):
    """Eyes on a Fibonacci sphere, widths/heights log-uniform in the given ranges."""
    g = torch.Generator().manual_seed(seed)
    cams = []
    for eye in fibonacci_eyes(n, radius):
        u = torch.rand(2, generator=g, dtype=torch.float64)
        W = int(round(math.exp(math.log(w[0]) + float(u[0]) * (math.log(w[1]) - math.log(w[0])))))
        H = int(round(math.exp(math.log(h[0]) + float(u[1]) * (math.log(h[1]) - math.log(h[0])))))
        cams.append((eye, W, H))
    return cams


def camera_rays(cams, device, fov_deg=60.0, target=(0.0, 0.0, 0.0)):
    """Pinhole rays of every camera, concatenated on the device: origins, unit directions, int32 view index.
    Pixel centres at +0.5, focal 0.5 W / tan(fov/2) on both axes, OpenCV axes (as scenes.pinhole_rays)."""
    from .scenes import look_at
    os_, ds_, vs_ = [], [], []
    for v, (eye, W, H) in enumerate(cams):
        Rcw, e = look_at(eye, target)
        f = 0.5 * W / math.tan(0.5 * math.radians(fov_deg))
        j, i = torch.meshgrid(torch.arange(H, device=device, dtype=torch.float64),
                              torch.arange(W, device=device, dtype=torch.float64), indexing="ij")
        dc = torch.stack([(i + 0.5 - 0.5 * W) / f, (j + 0.5 - 0.5 * H) / f, torch.ones_like(i)], -1).reshape(-1, 3)
        dw = dc @ Rcw.to(device).T
        dw = dw / dw.norm(dim=1, keepdim=True)
        ds_.append(dw.float())
        os_.append(e.to(device).float().expand(H * W, 3))
        vs_.append(torch.full((H * W,), v, dtype=torch.int32, device=device))
    return torch.cat(os_).contiguous(), torch.cat(ds_).contiguous(), torch.cat(vs_).contiguous()


class BatchStep:
    """LBVH build + forward + backward of one camera batch with fixed inputs; run eagerly or captured once into
    a HIP graph and replayed (outputs are then the graph's own tensors, rewritten by every replay)."""

    def __init__(self, params, cams, sh_degree=3, activation=1, kbuf=16, tile=0, tmin=0.0, tmax=100.0,
                 max_iters=500, grad_seed=1):
        self.params = [p.detach().contiguous() for p in params]
        self.device = self.params[0].device
        self.cams = cams
        self.eyes = torch.tensor([c[0] for c in cams], dtype=torch.float32).to(self.device).contiguous()
        self.rayo, self.rayd, self.view = camera_rays(cams, self.device)
        self.nr = self.rayo.shape[0]
        g = torch.Generator().manual_seed(grad_seed)
        self.grad = torch.randn((self.nr, 5), generator=g).to(self.device)
        self.sh_degree, self.activation = sh_degree, activation
        self.kbuf, self.tile = kbuf, tile
        self.tmin, self.tmax, self.max_iters = tmin, tmax, max_iters
        self.graph = None
        self.out = None

    def run(self):
        scene = T.Scene(*self.params, color_mode=1, sh_origin=self.cams[0][0], sh_degree=self.sh_degree,
                        activation=self.activation, view_origins=self.eyes).build()
        f = T.trace_forward(scene, self.rayo, self.rayd, self.tmin, self.tmax, self.max_iters, kbuf=self.kbuf,
                            traversal="wave", view_index=self.view, tile=self.tile, count_touches=False,
                            event_layout="slot")
        b = T.trace_backward(scene, self.rayo, self.rayd, self.tmin, self.tmax, self.max_iters, f, self.grad,
                             kbuf=self.kbuf)
        return {"color": f["color"], "iters": f["iters"], "initial_drgb": f["initial_drgb"], **b}

    def capture(self):
        """Capture build + forward + backward into one HIP graph (after an eager warm-up on a side stream)."""
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            self.run()
        torch.cuda.current_stream(self.device).wait_stream(s)
        torch.cuda.synchronize(self.device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.out = self.run()
        return self

    def replay(self):
        self.graph.replay()
        return self.out


def _cache_path():
    return os.environ.get("EVER_AUTOTUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "ever_amd",
                                                               "autotune.json"))


def workload_key(num_prims, num_rays, num_views, device):
    """Cache key: device name and the power-of-two buckets of the sizes."""
    name = torch.cuda.get_device_name(device)
    b = lambda x: int(2 ** round(math.log2(max(x, 1))))
    return f"{name}|prims~{b(num_prims)}|rays~{b(num_rays)}|views{num_views}"


def autotune(make_step, candidates=None, reps=3, key=None, use_cache=True):
    """Times `make_step(kbuf, tile).run()` (build + forward + backward) for each candidate (kbuf, tile) and returns
    (best, {candidate: ms}); with a key, results are cached in a JSON file and reused."""
    candidates = candidates or [(k, t) for k in (8, 16, 19, 32) for t in (0, 1, 2, 4)]
    path = _cache_path()
    cache = {}
    if key and use_cache and os.path.exists(path):
        try:
            with open(path) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
        if key in cache:
            return tuple(cache[key]["best"]), {tuple(map(int, k.split(","))): v for k, v in cache[key]["ms"].items()}
    times = {}
    for cand in candidates:
        step = make_step(*cand)
        step.run()  # warm-up (allocations, first-touch)
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            step.run()
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1))
        times[cand] = best
        del step
    best = min(times, key=times.get)
    if key:
        cache[key] = {"best": list(best), "ms": {f"{k[0]},{k[1]}": v for k, v in times.items()},
                      "when": time.strftime("%Y-%m-%d %H:%M:%S")}
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path, "w") as f:
                json.dump(cache, f, indent=1)
        except OSError:
            pass
    return best, times
