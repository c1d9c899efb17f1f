"""Learned termination proposal (no counterpart in the reference): a BF16 MLP on the MFMA units that predicts, per
ray, the depth at which the ray saturates, and the tracer's proposal-bounded first traversal that uses it.

The network (include/ever_amd.h, ever_mlp_*) maps a 63-dim encoding of (origin + tmin * direction, unit direction)
through 3 x 128 ReLU layers to y = log(depth); it is trained on the GPU with the library's own forward/backward
kernels (pinball loss at quantile tau, Adam) on (ray, saturation depth) pairs produced by the plain tracer. A
proposal only bounds where a ray first looks for events: a ray that has not terminated by its proposal continues
with the plain traversal, so the traced result is the plain one.
"""
import ctypes
import math

import torch

from . import _capi as C
from . import tracer as T

HID = 128
IN = 64
LOG_CUTOFF = 5.54        # the tracer's termination on log-transmittance (shaders.slang:191)
BACKGROUND_DEPTH = 64.0  # training target of rays that never saturate: beyond any scene of the workloads


def layout():
    """Offsets of the FP32 parameter block (same order as ever_mlp_param_count)."""
    off, o = {}, 0
    for name, size in (("w1", IN * HID), ("b1", HID), ("w2", HID * HID), ("b2", HID), ("w3", HID * HID),
                       ("b3", HID), ("w4", HID), ("b4", 1)):
        off[name] = (o, size)
        o += size
    return off, o


class ProposalNet:
    """Parameters, Adam state and workspaces of the proposal network on one device."""

    def __init__(self, device=None, seed=0):
        self.device = torch.device(device or "cuda")
        L = C.lib()
        off, count = layout()
        if count != L.ever_mlp_param_count():
            raise C.EverError("proposal network layout mismatch with the library")
        g = torch.Generator().manual_seed(seed)
        p = torch.zeros(count)
        for name, fan_in, fan_out in (("w1", 63, HID), ("w2", HID, HID), ("w3", HID, HID), ("w4", HID, 1)):
            o, size = off[name]
            w = torch.randn(size, generator=g) * math.sqrt(2.0 / fan_in)
            if name == "w1":
                w.view(IN, HID)[63:] = 0.0  # the zero feature column
            if name == "w4":
                w *= 0.1
            p[o:o + size] = w
        p[off["b4"][0]] = math.log(4.0)
        self.params = p.to(self.device)
        self.params_bf16 = torch.empty(count, dtype=torch.int16, device=self.device)
        self.adam_m = torch.zeros(count, device=self.device)
        self.adam_v = torch.zeros(count, device=self.device)
        self.step = 0
        self._ws = None
        with torch.cuda.device(self.device):
            C.check(L.ever_mlp_sync_bf16(C.ptr(self.params), C.ptr(self.params_bf16), C.stream_ptr()),
                    "ever_mlp_sync_bf16")

    def set_params(self, params):
        self.params.copy_(params)
        L = C.lib()
        with torch.cuda.device(self.device):
            C.check(L.ever_mlp_sync_bf16(C.ptr(self.params), C.ptr(self.params_bf16), C.stream_ptr()),
                    "ever_mlp_sync_bf16")

    def save(self, path):
        """Write the FP32 parameters, the Adam moments and the step count to a safetensors file (a checkpoint from
        which training resumes exactly)."""
        from safetensors.torch import save_file
        off, count = layout()
        save_file({"params": self.params.detach().cpu().contiguous(), "adam_m": self.adam_m.cpu().contiguous(),
                   "adam_v": self.adam_v.cpu().contiguous(),
                   "step": torch.tensor([self.step], dtype=torch.int64)},
                  path, metadata={"format": "ever_amd.ProposalNet", "hidden": str(HID), "inputs": str(IN),
                                  "param_count": str(count)})

    @classmethod
    def load(cls, path, device=None):
        """A network with the parameters, Adam state and step count of a file written by save()."""
        from safetensors import safe_open
        net = cls(device)
        with safe_open(path, framework="pt") as f:
            meta = f.metadata() or {}
            if meta.get("format") != "ever_amd.ProposalNet" or int(meta.get("param_count", -1)) != net.params.numel():
                raise ValueError(f"{path} is not a proposal network checkpoint of this layout")
            net.set_params(f.get_tensor("params").to(net.device))
            net.adam_m.copy_(f.get_tensor("adam_m"))
            net.adam_v.copy_(f.get_tensor("adam_v"))
            net.step = int(f.get_tensor("step")[0])
        return net

    def _workspace(self, nr):
        need = C.lib().ever_mlp_workspace_bytes(nr)
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need, dtype=torch.uint8, device=self.device)
        return self._ws

    def log_depth(self, rayo, rayd, tmin=0.0):
        nr = rayo.numel() // 3
        y = torch.empty(nr, dtype=torch.float32, device=self.device)
        self._infer(rayo, rayd, tmin, 1.0, y, None)
        return y

    def proposal(self, rayo, rayd, tmin=0.0, margin=1.05):
        """Proposed termination depth per ray: exp(y) * margin."""
        nr = rayo.numel() // 3
        out = torch.empty(nr, dtype=torch.float32, device=self.device)
        self._infer(rayo, rayd, tmin, margin, None, out)
        return out

    def _infer(self, rayo, rayd, tmin, margin, y, prop):
        nr = rayo.numel() // 3
        ws = self._workspace(nr)
        with torch.cuda.device(self.device):
            C.check(C.lib().ever_mlp_infer(C.ptr(self.params), C.ptr(self.params_bf16), C.ptr(rayo), C.ptr(rayd),
                                           nr, float(tmin), float(margin), C.ptr(y), C.ptr(prop), C.ptr(ws),
                                           ws.numel(), C.stream_ptr()), "ever_mlp_infer")

    def train_step(self, rayo, rayd, target, tmin=0.0, tau=0.7, lr=1e-3):
        """One Adam step on a batch; returns the batch loss (device scalar)."""
        self.step += 1
        return self._step(rayo, rayd, target, tmin, tau, lr, None)

    def gradient(self, rayo, rayd, target, tmin=0.0, tau=0.7):
        """Gradient of the batch loss w.r.t. the FP32 parameter block (no update)."""
        grad = torch.empty_like(self.params)
        self._step(rayo, rayd, target, tmin, tau, 0.0, grad)
        return grad

    def _step(self, rayo, rayd, target, tmin, tau, lr, grad):
        nr = rayo.numel() // 3
        ws = self._workspace(nr)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        with torch.cuda.device(self.device):
            C.check(C.lib().ever_mlp_train_step(C.ptr(self.params), C.ptr(self.params_bf16), C.ptr(self.adam_m),
                                                C.ptr(self.adam_v), C.ptr(rayo), C.ptr(rayd), C.ptr(target), nr,
                                                float(tmin), float(tau), float(lr), int(max(self.step, 1)),
                                                C.ptr(ws), ws.numel(), C.ptr(loss), C.ptr(grad), C.stream_ptr()),
                    "ever_mlp_train_step")
        return loss[0]


def saturation_depth(fwd):
    """Training target from a plain forward: the depth of the terminating event for rays that saturated
    (log-transmittance reached the tracer's cutoff), BACKGROUND_DEPTH for the others."""
    st = fwd["ray_state64"]
    sat = st[:, 4] >= LOG_CUTOFF
    return torch.where(sat, st[:, 5], torch.full_like(st[:, 5], BACKGROUND_DEPTH)).float().contiguous()


def random_cameras(n, radius=(3.0, 5.0), seed=4):
    """n eyes uniformly on the sphere with radius ~ U[radius]."""
    g = torch.Generator().manual_seed(seed)
    v = torch.randn((n, 3), generator=g, dtype=torch.float64)
    v = v / v.norm(dim=1, keepdim=True)
    r = radius[0] + (radius[1] - radius[0]) * torch.rand((n, 1), generator=g, dtype=torch.float64)
    return [tuple(float(x) for x in row) for row in (v * r)]


def make_pairs(scene, n_cameras=64, width=256, height=256, radius=(3.0, 5.0), seed=4, kbuf=None):
    """(origins, directions, targets) from plain traces of `scene` seen from random cameras at radius 3-5."""
    from .scenes import pinhole_rays
    os_, ds_, ts_ = [], [], []
    for eye in random_cameras(n_cameras, radius, seed):
        o, d = pinhole_rays(width, height, eye)
        o, d = o.to(scene.device).contiguous(), d.to(scene.device).contiguous()
        f = T.trace_forward(scene, o, d, 0.0, 1000.0, 500, store_events=False, kbuf=kbuf)
        os_.append(o)
        ds_.append(d)
        ts_.append(saturation_depth(f))
    return torch.cat(os_), torch.cat(ds_), torch.cat(ts_)


def fit(net, origins, directions, targets, epochs=3, batch=65536, lr=2e-3, tau=0.7, seed=5):
    """Adam over shuffled batches; returns the mean loss of the last epoch (host float)."""
    n = targets.numel()
    g = torch.Generator(device=origins.device).manual_seed(seed)
    last = 0.0
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g, device=origins.device)
        losses = []
        for b0 in range(0, n - batch + 1, batch):
            idx = perm[b0:b0 + batch]
            rate = lr * 0.5 * (1 + math.cos(math.pi * (ep * n + b0) / (epochs * n)))
            losses.append(net.train_step(origins[idx].contiguous(), directions[idx].contiguous(),
                                         targets[idx].contiguous(), 0.0, tau, rate))
        last = float(torch.stack(losses).mean()) if losses else 0.0
    return last
