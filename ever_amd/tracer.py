"""Host side of the HIP tracer: input validation, buffer allocation (PyTorch), C-ABI calls.

Mirrors the reference's pybind layer (splinetracers/fast_ellipsoid_splinetracer/py_binding.cpp):
input checks :33-61 and :104-116, output allocation :218-235, saved-for-backward tensors :151-193.
All buffers are PyTorch tensors on the tracer's device; every call is enqueued on the current stream and
nothing here synchronizes with the host.
"""
import ctypes
import os

import torch

from . import _capi as C

BUFFER_SIZE = 16  # events gathered per traversal (shaders.slang:18)
# k-buffer sizes with compiled kernels; results do not depend on the size, they follow the reference's 16-event
# rounds (which decide which exits beyond a cutting tmax are integrated) (19 is the largest whose per-wave LDS
# buffer, 19 x 64 x 8 B plus a 256 B prefetch area, still lets 16 waves share a CU's 160 KB)
KBUF_SIZES = (8, 16, 19, 32)
_KBUF = int(os.environ.get("EVER_KBUF", BUFFER_SIZE))


class StageTimer:
    """Records a pair of torch.cuda.Events around each library call (on the stream the kernels run on).

    Enabled by assigning an instance to ``tracer.stage_timer``; ``summary()`` returns mean milliseconds per
    stage after a synchronize.
    """

    def __init__(self):
        self.events = {}

    def start(self, name):
        e0 = torch.cuda.Event(enable_timing=True)
        e0.record()
        return name, e0

    def stop(self, token):
        name, e0 = token
        e1 = torch.cuda.Event(enable_timing=True)
        e1.record()
        self.events.setdefault(name, []).append((e0, e1))

    def summary(self):
        torch.cuda.synchronize()
        return {k: sum(a.elapsed_time(b) for a, b in v) / len(v) for k, v in self.events.items()}

    def reset(self):
        self.events = {}


stage_timer = None


def _t0(name):
    return stage_timer.start(name) if stage_timer is not None else None


def _t1(tok):
    if tok is not None:
        stage_timer.stop(tok)


def set_kbuf(k):
    """Select the k-buffer size (8, 16, 19 or 32) used by subsequent traces."""
    global _KBUF
    if k not in KBUF_SIZES:
        raise ValueError("kbuf must be 8, 16, 19 or 32")
    _KBUF = k


def get_kbuf():
    return _KBUF


# tree walk of the forward pass (EVER_TRAVERSAL_* in include/ever_amd.h); every mode gives the same results
TRAVERSALS = {"auto": 0, "ray": 1, "wave": 2}
# event-list layouts (EVER_EVENTS_BY_* in include/ever_amd.h)
EVENT_LAYOUTS = {"ray": C.EVER_EVENTS_BY_RAY, "slot": C.EVER_EVENTS_BY_SLOT}
_TRAVERSAL = os.environ.get("EVER_TRAVERSAL", "auto")


def set_traversal(mode):
    """Select the forward tree walk: "auto" (per-wave for sorted batches of >= 65536 rays), "ray" or "wave"."""
    global _TRAVERSAL
    if mode not in TRAVERSALS:
        raise ValueError(f"traversal must be one of {sorted(TRAVERSALS)}")
    _TRAVERSAL = mode


def get_traversal():
    return _TRAVERSAL


def _check_float(x, name, last=None, device=None):
    # py_binding.cpp:33-61 (CHECK_CUDA, CHECK_CONTIGUOUS, CHECK_FLOAT, last dimension)
    if not isinstance(x, torch.Tensor):
        raise TypeError(f"{name} must be a tensor")
    if not x.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if device is not None and x.device != device:
        raise RuntimeError(f"{name} must be on the same device")
    if not x.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")
    if x.dtype != torch.float32:
        raise RuntimeError(f"{name} must have float32 type")
    if last is not None and x.size(-1) != last:
        raise RuntimeError(f"{name} must have last dimension with size {last}")


class Scene:
    """Primitive set + LBVH workspace (replaces Primitives + GAS, py_binding.cpp:94-149)."""

    def __init__(self, means, scales, quats, densities, features, color_mode=C.EVER_COLOR_DC, sh_origin=None,
                 sh_degree=0, activation=0, view_origins=None):
        """view_origins: (V,3) origins of several views traced in one batch (colour mode SH); each ray then names
        its view (trace_forward view_index). Not in the reference, where each camera is a separate call."""
        device = means.device
        n = means.shape[0]
        _check_float(means, "means", 3, device)
        _check_float(scales, "scales", 3, device)
        _check_float(quats, "quats", 4, device)
        _check_float(features, "colors", 3, device)
        _check_float(densities, "densities", None, device)
        if scales.size(0) != n:
            raise RuntimeError("All inputs (scales) must have the same 0 dimension")
        if quats.size(0) != n:
            raise RuntimeError("All inputs (quats) must have the same 0 dimension")
        if features.size(0) != n:
            raise RuntimeError("All inputs (colors) must have the same 0 dimension")
        if densities.size(0) != n:
            raise RuntimeError("All inputs (densities) must have the same 0 dimension")
        if features.dim() != 3 or features.size(2) != 3:
            raise RuntimeError("Features must have 3 channels. (N, d, 3)")
        if densities.numel() != n:
            raise RuntimeError("densities must hold one value per primitive")
        self.device = device
        self.n = n
        self.tensors = (means, scales, quats, densities, features)
        self.feature_size = features.size(1)
        self.color_mode = color_mode
        self.sh_degree = sh_degree
        self.activation = activation
        self.sh_origin = [0.0, 0.0, 0.0] if sh_origin is None else [float(v) for v in sh_origin]
        self.num_views = 0
        self.view_origins = None
        self.view_colors = None
        if view_origins is not None:
            # a float32 tensor already on the device is used as is (no copy: usable under graph capture, where
            # sh_origin should then name view 0 so that nothing is read back)
            if torch.is_tensor(view_origins) and view_origins.device == device:
                vo = view_origins.reshape(-1, 3)
                _check_float(vo, "view_origins", 3, device)
            else:
                vo = torch.as_tensor(view_origins, dtype=torch.float32).reshape(-1, 3).to(device).contiguous()
            if color_mode != C.EVER_COLOR_SH:
                raise RuntimeError("several views need SH colours")
            self.num_views = vo.shape[0]
            self.view_origins = vo
            if sh_origin is None:
                self.sh_origin = [float(v) for v in vo[0].cpu()]
            if self.num_views > 1:
                self.view_colors = torch.empty((self.num_views, n, 4), dtype=torch.float32, device=device)
        self.prims = C.Prims(C.ptr(means), C.ptr(scales), C.ptr(quats), C.ptr(densities), C.ptr(features), n,
                             self.feature_size, color_mode, sh_degree, activation,
                             (ctypes.c_float * 3)(*self.sh_origin), self.num_views, C.ptr(self.view_origins),
                             C.ptr(self.view_colors))
        L = C.lib()
        self.ws_bytes = L.ever_scene_workspace_bytes(n)
        self.workspace = torch.empty(self.ws_bytes, dtype=torch.uint8, device=device)

    def build(self):
        L = C.lib()
        with torch.cuda.device(self.device):
            tok = _t0("build")
            C.check(L.ever_build_scene(ctypes.byref(self.prims), C.ptr(self.workspace), self.ws_bytes,
                                       C.stream_ptr()), "ever_build_scene")
            _t1(tok)
        return self

    def update_features(self, features, sh_origin=None, sh_degree=None):
        """New colour features for the built scene (Forward.reset_features, Forward.cpp:252-255): only the
        per-primitive colours are rewritten; the tree and geometry records stay as built. With SH colours the
        origin and degree of the SH evaluation may change too (a new camera for the same scene)."""
        _check_float(features, "colors", 3, self.device)
        if features.dim() != 3 or features.size(0) != self.n:
            raise RuntimeError("All inputs (colors) must have the same 0 dimension")
        means, scales, quats, densities, _ = self.tensors
        self.tensors = (means, scales, quats, densities, features)
        self.feature_size = features.size(1)
        self.prims.features = features.data_ptr()
        self.prims.feature_size = self.feature_size
        if sh_origin is not None:
            self.sh_origin = [float(v) for v in sh_origin]
            self.prims.sh_origin = (ctypes.c_float * 3)(*self.sh_origin)
        if sh_degree is not None:
            self.sh_degree = int(sh_degree)
            self.prims.sh_degree = self.sh_degree
        L = C.lib()
        with torch.cuda.device(self.device):
            tok = _t0("update_colors")
            C.check(L.ever_update_colors(ctypes.byref(self.prims), C.ptr(self.workspace), self.ws_bytes,
                                         C.stream_ptr()), "ever_update_colors")
            _t1(tok)
        return self


class SavedForBackward:
    """Mirror of fesSavedForBackward (py_binding.cpp:151-193)."""

    def __init__(self, states, diracs, faces, touch_count, iters):
        self.states = states
        self.diracs = diracs
        self.faces = faces
        self.touch_count = touch_count
        self.iters = iters


def trace_forward(scene, rayo, rayd, tmin, tmax, max_iters, store_events=True, zero_tail=False, kbuf=None,
                  traversal=None, proposal=None, view_index=None, tile=0, count_touches=True, event_layout="ray",
                  last_dirac=False):
    """Forward trace (Forward.trace_rays, py_binding.cpp:210-258). Returns a dict of tensors.

    count_touches=False skips the per-primitive event counts (touch_count is then None): one global atomic per
    event that only the caller's extras ever read. event_layout "ray" stores event k of ray r at
    tri_collection[r + k*R] (the reference's layout); "slot" stores it at the ray's position in ray_order, so that
    the rays traced together write, and the backward reads, contiguous rows (trace_backward follows the forward's
    choice). last_dirac=True also returns each ray's last integrated control point (the reference's
    SavedForBackward.diracs)."""
    if event_layout not in EVENT_LAYOUTS:
        raise ValueError("event_layout must be 'ray' or 'slot'")
    _check_float(rayo, "ray_origins", 3, scene.device)
    _check_float(rayd, "ray_directions", 3, scene.device)
    nr = rayo.numel() // 3
    if rayd.numel() != 3 * nr:
        raise RuntimeError("ray_origins and ray_directions must hold the same number of rays")
    if max_iters < 0:
        raise RuntimeError("max_iters must be >= 0")
    dev = scene.device
    L = C.lib()
    color = torch.empty((nr, 5), dtype=torch.float32, device=dev)
    states = torch.empty((nr, 16), dtype=torch.float32, device=dev)
    iters = torch.empty((nr,), dtype=torch.int32, device=dev)
    alloc = torch.zeros if zero_tail else torch.empty
    tri = alloc((nr * max_iters,), dtype=torch.int32, device=dev) if store_events else None
    touch = torch.empty((scene.n,), dtype=torch.int32, device=dev) if count_touches else None
    init = torch.empty((nr, 4), dtype=torch.float32, device=dev)
    st64 = torch.empty((nr, 8), dtype=torch.float64, device=dev)
    order = torch.empty((nr,), dtype=torch.int32, device=dev)
    dirac = torch.empty((nr, 4), dtype=torch.float32, device=dev) if last_dirac else None
    fws_bytes = L.ever_forward_workspace_bytes_for(nr, TRAVERSALS[traversal or _TRAVERSAL])
    fws = torch.empty(fws_bytes, dtype=torch.uint8, device=dev)
    if proposal is not None:
        _check_float(proposal, "proposal", None, dev)
        if proposal.numel() != nr:
            raise RuntimeError("proposal must hold one depth per ray")
    if scene.num_views > 1:
        if view_index is None or view_index.dtype != torch.int32 or view_index.numel() != nr:
            raise RuntimeError("view_index (int32, one per ray) is needed for a scene with several views")
    rays = C.Rays(C.ptr(rayo), C.ptr(rayd), nr, float(tmin), float(tmax), int(max_iters), kbuf or _KBUF,
                  TRAVERSALS[traversal or _TRAVERSAL], C.ptr(proposal), C.ptr(view_index), int(tile),
                  EVENT_LAYOUTS[event_layout])
    out = C.ForwardOut(C.ptr(color), C.ptr(states), C.ptr(iters), C.ptr(tri), C.ptr(touch), C.ptr(init),
                       C.ptr(st64), C.ptr(order), C.ptr(fws), fws_bytes, C.ptr(dirac))
    with torch.cuda.device(dev):
        tok = _t0("forward")
        C.check(L.ever_trace_forward(ctypes.byref(scene.prims), C.ptr(scene.workspace), scene.ws_bytes,
                                     ctypes.byref(rays), ctypes.byref(out), C.stream_ptr()), "ever_trace_forward")
        _t1(tok)
    return dict(color=color, states=states, iters=iters, tri_collection=tri, touch_count=touch, initial_drgb=init,
                ray_state64=st64, ray_order=order, rays=rays, view_index=view_index, event_layout=event_layout,
                last_dirac=dirac, proposal=proposal)


# rays per chunk of the recomputing adjoint: its event list holds chunk x max_iters ids (512 MB at 500 events)
RECOMPUTE_CHUNK = 262144


def trace_backward(scene, rayo, rayd, tmin, tmax, max_iters, fwd, grad_out, wcts=None, kbuf=None,
                   chunk_rays=None):
    """Adjoint pass (SplineTracer.backward, fast_ellipsoid_splinetracer.py:105-219). Returns a dict of grads.

    When the forward kept no event list (fwd["tri_collection"] is None) the rays are traced again, chunk_rays
    (default RECOMPUTE_CHUNK) at a time, into one chunk-sized list (ever_trace_backward_recompute); the
    proposal, view index and traversal of fwd["rays"] are reused."""
    dev = scene.device
    n = scene.n
    nr = rayo.numel() // 3
    L = C.lib()
    means, scales, quats, densities, features = scene.tensors
    g = grad_out.contiguous().float()
    out = dict(
        dmeans=torch.empty((n, 3), dtype=torch.float32, device=dev),
        dscales=torch.empty((n, 3), dtype=torch.float32, device=dev),
        dquats=torch.empty((n, 4), dtype=torch.float32, device=dev),
        ddensities=torch.empty((n,), dtype=torch.float32, device=dev),
        dfeatures=torch.empty(features.shape, dtype=torch.float32, device=dev),
        drayo=torch.empty((nr, 3), dtype=torch.float32, device=dev),
        drayd=torch.empty((nr, 3), dtype=torch.float32, device=dev),
        dinitial_drgb=torch.empty((nr, 4), dtype=torch.float32, device=dev),
    )
    nw = 0
    if wcts is not None:
        wcts = wcts.contiguous().float()
        _check_float(wcts, "wcts", 4, dev)
        nw = wcts.numel() // 16
        out["dmeans2D"] = torch.empty((n, 2), dtype=torch.float32, device=dev)
    vi = fwd.get("view_index")
    recompute = fwd.get("tri_collection") is None
    frays = fwd.get("rays")
    traversal = frays.traversal if frays is not None else TRAVERSALS[_TRAVERSAL]
    chunk = int(chunk_rays or RECOMPUTE_CHUNK)
    if recompute:
        ws_bytes = L.ever_backward_recompute_workspace_bytes(n, nr, nw, int(max_iters), traversal, chunk)
    else:
        ws_bytes = L.ever_backward_workspace_bytes(n, nr, nw)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    rays = C.Rays(C.ptr(rayo), C.ptr(rayd), nr, float(tmin), float(tmax), int(max_iters), kbuf or _KBUF, traversal,
                  C.ptr(fwd.get("proposal")) if recompute else None, C.ptr(vi), 0,
                  EVENT_LAYOUTS[fwd.get("event_layout", "ray")])
    vgrad = None
    if scene.num_views > 1:
        vgrad = torch.empty((scene.num_views, n, 4), dtype=torch.float32, device=dev)
    io = C.BackwardIO(C.ptr(g), C.ptr(fwd["iters"]), C.ptr(fwd["tri_collection"]), C.ptr(fwd["states"]),
                      C.ptr(fwd["ray_state64"]), C.ptr(fwd.get("ray_order")), C.ptr(wcts), nw, C.ptr(out["dmeans"]),
                      C.ptr(out["dscales"]),
                      C.ptr(out["dquats"]), C.ptr(out["ddensities"]), C.ptr(out["dfeatures"]), C.ptr(out["drayo"]),
                      C.ptr(out["drayd"]), C.ptr(out.get("dmeans2D")), C.ptr(out["dinitial_drgb"]), C.ptr(ws),
                      ws_bytes, C.ptr(vgrad))
    with torch.cuda.device(dev):
        tok = _t0("backward")
        if recompute:
            C.check(L.ever_trace_backward_recompute(ctypes.byref(scene.prims), C.ptr(scene.workspace),
                                                    scene.ws_bytes, ctypes.byref(rays), ctypes.byref(io), chunk,
                                                    C.stream_ptr()), "ever_trace_backward_recompute")
        else:
            C.check(L.ever_trace_backward(ctypes.byref(scene.prims), C.ptr(scene.workspace), scene.ws_bytes,
                                          ctypes.byref(rays), ctypes.byref(io), C.stream_ptr()), "ever_trace_backward")
        _t1(tok)
    return out
