"""Drop-in replacement of the reference's splinetracers/fast_ellipsoid_splinetracer.py on MI355X.

Same names, arguments, defaults, output layout and extras as the reference
(splinetracers/fast_ellipsoid_splinetracer.py:37-256): ``trace_rays`` returns an (R, 5) float32 tensor
[R, G, B, opacity, distortion] through the autograd Function ``SplineTracer``; with ``return_extras=True`` it
also returns dict(tri_collection, iters, opacity, touch_count, distortion_loss, saved).

Everything runs in the HIP library: the LBVH is rebuilt on every call (replacing the reference's OptiX GAS,
:59-71), the forward traces with a per-ray k-buffer, the backward replays each ray's stored events.

Beyond the reference: ``trace_rays_sh`` fuses the caller's SH prepass
(integration_utils/gaussian_renderer/ever.py:178-180: eval_sh at the camera, softplus(beta=10), RGB2SH) into
the tracer; its result equals eval_sh -> softplus -> RGB2SH -> trace_rays.
"""
from typing import Any

import torch
from torch.autograd import Function

from .. import _capi as C
from .. import tracer as T
from . import sp  # noqa: F401  (the reference's compiled module of the same name, for its inference callers)

# gradient clamps of the reference's backward (fast_ellipsoid_splinetracer.py:201-211) are applied in the
# library's per-primitive epilogue and in the per-ray adjoint kernel.

# Adjoint event source. "stored": the forward keeps every ray's event list (R * max_iters int32, the reference's
# tri_collection) for the backward. "recompute": the forward keeps none; the backward traces the rays again, a
# chunk at a time (same scene workspace, same deterministic walk, hence the same events), into one list of
# chunk * max_iters ids, so no list of size R * max_iters exists at any time. Gradients are equal up to the order
# of float atomic additions; "recompute" costs one more forward.
_EVENT_MODE = "stored"
_CHUNK_RAYS = T.RECOMPUTE_CHUNK


def set_event_mode(mode, chunk_rays=None):
    """Select "stored" or "recompute" event lists; chunk_rays sets the rays traced per chunk of the recomputing
    adjoint (its event list holds chunk_rays * max_iters ids)."""
    global _EVENT_MODE, _CHUNK_RAYS
    if mode not in ("stored", "recompute"):
        raise ValueError("event mode must be 'stored' or 'recompute'")
    if chunk_rays is not None:
        if int(chunk_rays) < 1:
            raise ValueError("chunk_rays must be >= 1")
        _CHUNK_RAYS = int(chunk_rays)
    _EVENT_MODE = mode


def get_event_mode():
    return _EVENT_MODE


def _run_forward(ctx, mean, scale, quat, density, color, rayo, rayd, tmin, tmax, max_prim_size, wcts, max_iters,
                 return_extras, color_mode, sh_origin, sh_degree, activation, proposal=None):
    ctx.device = rayo.device
    if mean.device != ctx.device:
        raise RuntimeError("mean must be on the same device as the rays")
    mean = mean.contiguous()
    scale = scale.contiguous()
    density = density.contiguous()
    quat = quat.contiguous()
    color = color.contiguous()
    scene = T.Scene(mean, scale, quat, density, color, color_mode=color_mode, sh_origin=sh_origin,
                    sh_degree=sh_degree, activation=activation).build()
    recompute = _EVENT_MODE == "recompute" and not return_extras
    # The forward's per-primitive touch counts reach the caller only through the extras (the reference's backward
    # counts its own, fast_ellipsoid_splinetracer.py:137), so they are only counted when extras are asked for.
    # Event lists the caller never sees are stored by processing slot (contiguous rows per wave); the extras'
    # tri_collection keeps the reference's by-ray layout.
    layout = "ray" if return_extras else "slot"
    out = T.trace_forward(scene, rayo, rayd, tmin, tmax, max_iters, store_events=not recompute,
                          zero_tail=return_extras, proposal=proposal, count_touches=return_extras,
                          event_layout=layout, last_dirac=return_extras)
    ctx.event_layout = layout
    ctx.recompute = recompute
    ctx.proposal = proposal
    ctx.rays = out["rays"]
    ctx.chunk_rays = _CHUNK_RAYS
    ctx.scene = scene
    ctx.max_iters = max_iters
    ctx.max_prim_size = max_prim_size
    ctx.tmin = tmin
    ctx.tmax = tmax
    ctx.kbuf = T.get_kbuf()
    ctx.save_for_backward(mean, scale, quat, density, color, rayo, rayd, out["tri_collection"], wcts,
                          out["initial_drgb"], out["states"], out["iters"], out["ray_state64"], out["ray_order"],
                          scene.workspace)
    color_and_loss = out["color"]
    if return_extras:
        nr = rayo.shape[0]
        saved = T.SavedForBackward(
            states=out["states"],
            diracs=out["last_dirac"],
            # last_tri, never updated by the reference (shaders.slang:187, 233)
            faces=torch.full((nr,), -1, dtype=torch.int32, device=ctx.device),
            touch_count=out["touch_count"],
            iters=out["iters"],
        )
        return color_and_loss, dict(
            tri_collection=out["tri_collection"],
            iters=out["iters"],
            opacity=color_and_loss[:, 3],
            touch_count=out["touch_count"],
            distortion_loss=color_and_loss[:, 4],
            saved=saved,
        )
    return color_and_loss


def _run_backward(ctx, grad_output):
    (mean, scale, quat, density, features, rayo, rayd, tri_collection, wcts, initial_drgb, states, iters, st64,
     order, _ws) = ctx.saved_tensors
    # recompute: no event list was kept (tri_collection is None) and the adjoint traces the rays again in chunks
    fwd = dict(tri_collection=tri_collection, states=states, iters=iters, ray_state64=st64,
               initial_drgb=initial_drgb, ray_order=order, event_layout=ctx.event_layout, rays=ctx.rays,
               proposal=ctx.proposal)
    g = T.trace_backward(ctx.scene, rayo, rayd, ctx.tmin, ctx.tmax, ctx.max_iters, fwd, grad_output, wcts=wcts,
                         kbuf=ctx.kbuf, chunk_rays=ctx.chunk_rays)
    return (g["dmeans"], g["dscales"], g["dquats"], g["ddensities"].reshape(density.shape), g["dfeatures"],
            g["drayo"], g["drayd"], g.get("dmeans2D"))


class SplineTracer(Function):
    """Autograd Function with the reference's 14 arguments (fast_ellipsoid_splinetracer.py:37-219)."""

    @staticmethod
    def forward(ctx: Any, mean: torch.Tensor, scale: torch.Tensor, quat: torch.Tensor, density: torch.Tensor,
                color: torch.Tensor, rayo: torch.Tensor, rayd: torch.Tensor, tmin: float, tmax: float,
                max_prim_size: float, mean2D: torch.Tensor, wcts: torch.Tensor, max_iters: int,
                return_extras: bool = False):
        return _run_forward(ctx, mean, scale, quat, density, color, rayo, rayd, tmin, tmax, max_prim_size, wcts,
                            max_iters, return_extras, C.EVER_COLOR_DC, None, 0, 0)

    @staticmethod
    def backward(ctx, grad_output: torch.Tensor, *unused):
        dm, ds, dq, dd, df, dro, drd, dm2 = _run_backward(ctx, grad_output)
        return (dm, ds, dq, dd, df, dro, drd, None, None, None, dm2, None, None, None)


class SplineTracerSH(Function):
    """SplineTracer with the per-primitive SH colour prepass fused into the scene build."""

    @staticmethod
    def forward(ctx: Any, mean, scale, quat, density, sh_features, rayo, rayd, tmin, tmax, max_prim_size, mean2D,
                wcts, max_iters, return_extras, sh_origin, sh_degree, activation, proposal=None):
        return _run_forward(ctx, mean, scale, quat, density, sh_features, rayo, rayd, tmin, tmax, max_prim_size,
                            wcts, max_iters, return_extras, C.EVER_COLOR_SH, sh_origin, sh_degree, activation,
                            proposal)

    @staticmethod
    def backward(ctx, grad_output: torch.Tensor, *unused):
        dm, ds, dq, dd, df, dro, drd, dm2 = _run_backward(ctx, grad_output)
        return (dm, ds, dq, dd, df, dro, drd, None, None, None, dm2, None, None, None, None, None, None, None)


def trace_rays(
    mean: torch.Tensor,
    scale: torch.Tensor,
    quat: torch.Tensor,
    density: torch.Tensor,
    features: torch.Tensor,
    rayo: torch.Tensor,
    rayd: torch.Tensor,
    tmin: float = 0.0,
    tmax: float = 1000,
    max_prim_size: float = 3,
    dL_dmeans2D=None,
    wcts=None,
    max_iters: int = 500,
    return_extras: bool = False,
):
    """fast_ellipsoid_splinetracer.py:222-254"""
    return SplineTracer.apply(mean, scale, quat, density, features, rayo, rayd, tmin, tmax, max_prim_size,
                              dL_dmeans2D, wcts, max_iters, return_extras)


trace_rays.uses_density = True


def trace_rays_sh(mean, scale, quat, density, sh_features, rayo, rayd, sh_origin, sh_degree, tmin=0.0, tmax=1000,
                  max_prim_size=3, dL_dmeans2D=None, wcts=None, max_iters=500, return_extras=False,
                  activation="softplus", proposal=None):
    """trace_rays with colour = [softplus_10](SH(sh_features, normalize(mean - sh_origin))).

    Equals the reference caller's eval_sh -> softplus(beta=10) -> RGB2SH -> trace_rays
    (integration_utils/gaussian_renderer/ever.py:178-204); gradients flow to sh_features only through the
    colour (no gradient to the means through the SH direction, as eval_sh.py:68). `proposal` (optional, (R,)
    float32) bounds each ray's first traversal (ever_amd.proposal); the result is the same with or without it.
    """
    if isinstance(sh_origin, torch.Tensor):
        sh_origin = [float(v) for v in sh_origin.detach().reshape(3).cpu()]
    act = {None: 0, "none": 0, "softplus": 1}[activation]
    return SplineTracerSH.apply(mean, scale, quat, density, sh_features, rayo, rayd, tmin, tmax, max_prim_size,
                                dL_dmeans2D, wcts, max_iters, return_extras, tuple(sh_origin), int(sh_degree), act,
                                proposal)


trace_rays_sh.uses_density = True
