"""The reference's compiled module ``sp`` (splinetracers/fast_ellipsoid_splinetracer/py_binding.cpp:261-282) on the
HIP tracer: the same classes, methods, arguments, checks and output dict, so that code written against it (the
reference's integration_utils/gaussian_renderer/fast_renderer.py:50-63 and :112-119) runs unchanged.

    ctx  = sp.OptixContext(device)                       # no state: the HIP library has no context object
    prims = sp.Primitives(device)
    prims.add_primitives(means, scales, quats, half_attribs, densities, colors)
    gas  = sp.GAS(ctx, device, prims, enable_anyhit, fast_build, enable_rebuild)   # LBVH build
    fwd  = sp.Forward(ctx, device, prims, enable_backward)
    prims.set_features(new_colors); fwd.update_model(prims)                        # colours only
    out  = fwd.trace_rays(gas, rayo, rayd, tmin, tmax, max_iters, max_prim_size)

Colours follow the reference's forward shader: SH_C0 * colors[i, 0, :] + 0.5 (shaders.slang:65-82). As in the
reference, the GAS holds the geometry as it was when the GAS was built; colours are read when the rays are traced,
from the features the Forward was given last (at construction or by update_model), so ``trace_rays`` rewrites the
scene's per-primitive colours from them before every trace (one pass over the primitives, no tree rebuild).
A GAS traced by Forwards of different features is therefore correct for each of them.
"""
import torch

from .. import _capi as C
from .. import tracer as T


class OptixContext:
    """py_binding.cpp:70-92. Holds the device; loading the HIP library here fails loudly if it is not built."""

    def __init__(self, device):
        self.device = torch.device(device)
        C.lib()


class SavedForBackward:
    """py_binding.cpp:151-193 (read-only properties states, diracs, faces, touch_count, iters)."""

    def __init__(self, states, diracs, faces, touch_count, iters):
        self._t = dict(states=states, diracs=diracs, faces=faces, touch_count=touch_count, iters=iters)

    states = property(lambda self: self._t["states"])
    diracs = property(lambda self: self._t["diracs"])
    faces = property(lambda self: self._t["faces"])
    touch_count = property(lambda self: self._t["touch_count"])
    iters = property(lambda self: self._t["iters"])


class Primitives:
    """py_binding.cpp:94-139."""

    def __init__(self, device):
        self.device = torch.device(device)
        self.tensors = None
        self.features = None
        self.num_prims = 0

    def add_primitives(self, means, scales, quats, half_attribs, densities, colors):
        """py_binding.cpp:99-132: the same checks and messages. half_attribs (a float16 copy of means, scales and
        quats that only the reference's OptiX intersection program reads) is checked for its type and not used."""
        n = means.size(0)
        T._check_float(means, "means", 3)
        T._check_float(scales, "scales", 3)
        T._check_float(quats, "quats", 4)
        T._check_float(colors, "colors", 3)
        if scales.size(0) != n:
            raise RuntimeError("All inputs (scales) must have the same 0 dimension")
        if quats.size(0) != n:
            raise RuntimeError("All inputs (quats) must have the same 0 dimension")
        if colors.size(0) != n:
            raise RuntimeError("All inputs (colors) must have the same 0 dimension")
        if densities.size(0) != n:
            raise RuntimeError("All inputs (densities) must have the same 0 dimension")
        if colors.dim() != 3 or colors.size(2) != 3:
            raise RuntimeError("Features must have 3 channels. (N, d, 3)")
        if not isinstance(half_attribs, torch.Tensor) or half_attribs.dtype != torch.float16:
            raise RuntimeError("half_attribs: expected scalar type Half")
        self.tensors = (means, scales, quats, densities)
        self.features = colors
        self.num_prims = n

    def set_features(self, colors):
        """py_binding.cpp:133-138."""
        T._check_float(colors, "colors", 3)
        if colors.size(0) != self.num_prims:
            raise RuntimeError("All inputs (colors) must have the same 0 dimension")
        self.features = colors


def _require(prims):
    if prims.tensors is None:
        raise RuntimeError("Primitives has no primitives: call add_primitives first")


class GAS:
    """py_binding.cpp:141-149 (GAS.cpp:56-153): builds the LBVH of the primitives. enable_anyhit, fast_build and
    enable_rebuild select OptiX build options and have no counterpart here (one build type)."""

    def __init__(self, context, device, prims, enable_anyhit, fast_build, enable_rebuild):
        _require(prims)
        means, scales, quats, densities = prims.tensors
        self.device = torch.device(device)
        self.scene = T.Scene(means, scales, quats, densities, prims.features, color_mode=C.EVER_COLOR_DC).build()
        lay = C.scene_layout(self.scene.n)
        ws32 = self.scene.workspace.view(torch.int32)
        self._count = ws32[lay["header"] // 4 + 8: lay["header"] // 4 + 9]
        self._cand = ws32[lay["cand"] // 4: lay["cand"] // 4 + self.scene.n]


class Forward:
    """py_binding.cpp:195-259 (Forward.cpp:84-255)."""

    def __init__(self, context, device, prims, enable_backward):
        _require(prims)
        self.device = torch.device(device)
        self.enable_backward = bool(enable_backward)
        self.num_prims = prims.num_prims
        self.features = prims.features

    def update_model(self, prims):
        """py_binding.cpp:207-209 (Forward::reset_features, Forward.cpp:252-255): the features of the next traces."""
        self.features = prims.features

    def trace_rays(self, gas, ray_origins, ray_directions, tmin, tmax, max_iters, max_prim_size):
        """py_binding.cpp:210-258. Returns dict(color (R,4), saved, tri_collection (R*max_iters,) int32,
        initial_drgb (R,4), initial_touch_inds (N,) int32, initial_touch_count (1,) int32).

        With enable_backward False the event list and the per-primitive touch counts are not written (the
        reference's fast_shaders.slang:220-221 drops both writes) and are zero, as the reference's buffers; the
        event list is then a read-only zero-stride view.
        initial_touch_inds holds, in its first initial_touch_count entries, the primitives whose bounds lie within
        tmin of the first ray's origin (initialize_density.cu:69-105), in no particular order; the rest are zero.
        max_prim_size only bounds the reference's OptiX primitive size and is not used."""
        T._check_float(ray_origins, "ray_origins", 3)
        T._check_float(ray_directions, "ray_directions", 3)
        scene = gas.scene
        if self.num_prims != scene.n:
            raise RuntimeError("Forward and GAS were made from different primitive sets")
        with torch.no_grad():
            scene.update_features(self.features)
            nr = ray_origins.numel() // 3
            out = T.trace_forward(scene, ray_origins, ray_directions, tmin, tmax, max_iters,
                                  store_events=self.enable_backward, zero_tail=self.enable_backward,
                                  count_touches=self.enable_backward, event_layout="ray", last_dirac=True)
            dev = scene.device
            tri = out["tri_collection"]
            touch = out["touch_count"]
            if tri is None:
                # all zeros, as a zero-stride view: R * max_iters int32 of HBM are not filled for nothing
                tri = torch.zeros((1,), dtype=torch.int32, device=dev).expand(nr * max_iters)
                touch = torch.zeros((scene.n,), dtype=torch.int32, device=dev)
            count = gas._count.clone()
            inds = torch.where(torch.arange(scene.n, device=dev, dtype=torch.int32) < count, gas._cand,
                               torch.zeros((), dtype=torch.int32, device=dev))
            saved = SavedForBackward(
                states=out["states"],
                diracs=out["last_dirac"],
                faces=torch.full((nr,), -1, dtype=torch.int32, device=dev),  # never updated (shaders.slang:187, 233)
                touch_count=touch,
                iters=out["iters"],
            )
            return dict(color=out["color"][:, :4].contiguous(), saved=saved, tri_collection=tri,
                        initial_drgb=out["initial_drgb"], initial_touch_inds=inds, initial_touch_count=count)
