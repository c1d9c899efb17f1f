"""The C-ABI library loads, exports every function include/ever_amd.h declares, and validates arguments on the
host (no GPU work is issued by these calls)."""
import ctypes
import os
import re

import pytest

from ever_amd import _capi as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "ever_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(ever_\w+)\s*\(", text, flags=re.M)))


def test_every_header_function_is_exported_and_bound():
    names = header_functions()
    assert len(names) >= 12
    L = C.lib()
    for n in names:
        assert hasattr(L, n), n
        assert n in C.SIGNATURES, n


def test_abi_version_and_struct_sizes():
    L = C.lib()
    assert L.ever_abi_version() == 9
    sizes = (ctypes.c_int64 * 4)()
    assert L.ever_struct_sizes(sizes) == 0
    assert list(sizes) == [ctypes.sizeof(C.Prims), ctypes.sizeof(C.Rays), ctypes.sizeof(C.ForwardOut),
                           ctypes.sizeof(C.BackwardIO)]


def test_workspace_queries_grow_with_size():
    L = C.lib()
    a, b = L.ever_scene_workspace_bytes(1000), L.ever_scene_workspace_bytes(100000)
    assert 0 < a < b
    assert L.ever_scene_workspace_bytes(100000) >= 100000 * 64 * 2
    assert L.ever_forward_workspace_bytes(10) > 0
    # the per-ray walk spills up to 160 stack entries per ray, the per-wave walk none; AUTO takes the per-ray walk
    # below 65536 rays
    for nr in (10, 70000):
        ray, wave = L.ever_forward_workspace_bytes_for(nr, 1), L.ever_forward_workspace_bytes_for(nr, 2)
        assert ray - wave >= 4 * 150 * nr
        assert L.ever_forward_workspace_bytes(nr) == (ray if nr < 65536 else wave)
    # the recomputing adjoint holds one chunk's event list (chunk x max_iters ids), not the call's
    n, nr, it = 100000, 2_000_000, 500
    stored = L.ever_backward_workspace_bytes(n, nr, 0) + 4 * nr * it
    chunk = L.ever_backward_recompute_workspace_bytes(n, nr, 0, it, 0, 131072)
    assert chunk < L.ever_backward_recompute_workspace_bytes(n, nr, 0, it, 0, 262144) < stored - 3e9
    assert L.ever_backward_workspace_bytes(n, nr, 0) + 4 * 131072 * it <= chunk < stored - 3.5e9
    assert L.ever_backward_recompute_workspace_bytes(n, nr, 0, it, 0, 0) == 0
    lay = C.scene_layout(1000)
    offs = [lay[k] for k in C.LAYOUT_NAMES]
    assert offs == sorted(offs) and lay["total"] == a
    assert all(o % 256 == 0 for o in offs)


def test_invalid_arguments_return_status_codes():
    L = C.lib()
    assert L.ever_build_scene(None, None, 0, None) == 1
    assert b"prims is null" in L.ever_last_error()
    p = C.Prims()
    p.num_prims = 10
    p.feature_size = 1
    assert L.ever_build_scene(ctypes.byref(p), None, 0, None) == 1  # null arrays
    p.num_prims = 0
    assert L.ever_build_scene(ctypes.byref(p), None, 0, None) == 1  # null workspace
    buf = ctypes.create_string_buffer(16)
    assert L.ever_build_scene(ctypes.byref(p), ctypes.cast(buf, ctypes.c_void_p), 0, None) == 3
    assert L.ever_update_colors(None, None, 0, None) == 1
    assert L.ever_update_colors(ctypes.byref(p), None, 0, None) == 1  # null workspace
    assert L.ever_update_colors(ctypes.byref(p), ctypes.cast(buf, ctypes.c_void_p), 0, None) == 3
    r = C.Rays()
    r.kbuf = 7
    assert L.ever_trace_forward(ctypes.byref(p), ctypes.cast(buf, ctypes.c_void_p), 16, ctypes.byref(r), None,
                                None) == 1
    assert b"kbuf" in L.ever_last_error()
    r.kbuf = 16
    r.traversal = 3
    assert L.ever_trace_forward(ctypes.byref(p), ctypes.cast(buf, ctypes.c_void_p), 16, ctypes.byref(r), None,
                                None) == 1
    assert b"traversal" in L.ever_last_error()
    r.traversal = 0
    r.event_layout = 2
    assert L.ever_trace_forward(ctypes.byref(p), ctypes.cast(buf, ctypes.c_void_p), 16, ctypes.byref(r), None,
                                None) == 1
    assert b"event_layout" in L.ever_last_error()
    # by-slot event lists cannot be read back without the forward's processing order
    r.event_layout = C.EVER_EVENTS_BY_SLOT
    r.num_rays = 1
    r.origins = r.directions = 16  # never dereferenced: validation fails first
    io = C.BackwardIO()
    assert L.ever_trace_backward(ctypes.byref(p), ctypes.cast(buf, ctypes.c_void_p), 16, ctypes.byref(r),
                                 ctypes.byref(io), None) == 1
    assert b"ray_order" in L.ever_last_error()
    assert L.ever_eval_sh(None, None, None, 5, 1, 4, None, None) == 1
    lay = C.PlyLayout()
    assert L.ever_scene_from_records(None, 1, None, None, None, None, None, None, None, None) == 1
    lay.num_columns, lay.num_rest = 10, 4
    assert L.ever_scene_from_records(None, 1, ctypes.byref(lay), None, None, None, None, None, None, None) == 1
    assert b"num_rest" in L.ever_last_error()
    lay.num_rest, lay.f_dc[0] = 0, 12
    assert L.ever_scene_from_records(None, 1, ctypes.byref(lay), None, None, None, None, None, None, None) == 1
    assert b"out of range" in L.ever_last_error()
    assert L.ever_morton_order(None, 5, None, None, 0, None) == 1
    assert L.ever_morton_order(ctypes.cast(buf, ctypes.c_void_p), 5, ctypes.cast(buf, ctypes.c_void_p),
                               ctypes.cast(buf, ctypes.c_void_p), 16, None) == 3
    assert L.ever_morton_workspace_bytes(1000) > L.ever_morton_workspace_bytes(10) > 0
    # the variant without a workspace: argument checks before any GPU call, and its size limit
    assert L.ever_morton_order_noalloc(None, 5, None, None) == 1
    big = C.EVER_MORTON_NOALLOC_MAX_PRIMS + 1
    assert L.ever_morton_order_noalloc(ctypes.cast(buf, ctypes.c_void_p), big, ctypes.cast(buf, ctypes.c_void_p),
                                       None) == 1
    assert b"EVER_MORTON_NOALLOC_MAX_PRIMS" in L.ever_last_error()
    assert L.ever_morton_order_noalloc(None, 0, None, None) == 0
    # the recomputing adjoint checks its arguments as ever_trace_backward does, without an event list
    assert L.ever_trace_backward_recompute(None, None, 0, None, None, 1024, None) == 1
    assert b"prims is null" in L.ever_last_error()
    cam = C.Camera()
    assert L.ever_camera_rays(None, None, None, None, None) == 1
    cam.width, cam.height, cam.fx, cam.fy = 4, 3, 1.0, 1.0
    cam.model = 3
    assert L.ever_camera_rays(ctypes.byref(cam), None, None, None, None) == 1
    assert b"model" in L.ever_last_error()
    cam.model, cam.fx = C.EVER_CAMERA_FISHEYE, 0.0
    assert L.ever_camera_rays(ctypes.byref(cam), None, None, None, None) == 1
    assert b"focal" in L.ever_last_error()
    cam.fx = 1.0
    assert L.ever_camera_rays(ctypes.byref(cam), None, None, None, None) == 1  # null c2w / outputs
    cam.width = 0
    assert L.ever_camera_rays(ctypes.byref(cam), None, None, None, None) == 0  # empty image: nothing to do


def test_python_layer_validates_like_the_reference():
    import torch
    from ever_amd import tracer as T
    with pytest.raises(RuntimeError, match="CUDA"):
        T.Scene(torch.zeros(2, 3), torch.zeros(2, 3), torch.zeros(2, 4), torch.zeros(2), torch.zeros(2, 1, 3))


def test_sp_module_validates_like_the_reference():
    # py_binding.cpp:99-138: tensors are checked before anything is built
    import torch
    from ever_amd.splinetracers.fast_ellipsoid_splinetracer import sp
    sp.OptixContext("cpu")  # loads the library
    prims = sp.Primitives("cpu")
    z = lambda *s: torch.zeros(*s)
    with pytest.raises(RuntimeError, match="means must be a CUDA tensor"):
        prims.add_primitives(z(2, 3), z(2, 3), z(2, 4), z(2, 10).half(), z(2), z(2, 1, 3))
    with pytest.raises(RuntimeError, match="call add_primitives first"):
        sp.GAS(None, "cpu", prims, True, False, True)
    with pytest.raises(RuntimeError, match="call add_primitives first"):
        sp.Forward(None, "cpu", prims, False)


def test_camera_module_validates():
    import torch
    from ever_amd import cameras as CAM
    with pytest.raises(RuntimeError, match="CUDA"):
        CAM.camera_rays(4, 4, 1.0, 1.0, torch.eye(4))
    with pytest.raises(ValueError, match="projection"):
        CAM.projection_code("orthographic")
    assert CAM.projection_code(None) == 0 and CAM.projection_code("PANO") == 2
