"""ctypes binding of the C ABI in include/ever_amd.h (libever_amd.so, built for gfx950).

PyTorch provides device memory and streams; every compute call goes to the HIP library. There is no fallback:
if the library is missing or fails to load, `lib()` raises.
"""
import ctypes
import os

import torch  # noqa: F401  (loads PyTorch's HIP runtime first, so the library binds to the same runtime)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("EVER_AMD_LIB", os.path.join(_HERE, "libever_amd.so"))

EVER_OK = 0
EVER_ERR_INVALID = 1
EVER_ERR_HIP = 2
EVER_ERR_WORKSPACE = 3
EVER_MORTON_NOALLOC_MAX_PRIMS = 1048576
EVER_COLOR_DC = 0
EVER_COLOR_SH = 1
EVER_EVENTS_BY_RAY = 0
EVER_EVENTS_BY_SLOT = 1
EVER_CAMERA_PERSPECTIVE = 0
EVER_CAMERA_FISHEYE = 1
EVER_CAMERA_PANORAMIC = 2
EVER_PLY_ACTIVATE_NONE = 0
EVER_PLY_ACTIVATE_3DGS = 1

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32


class Prims(ctypes.Structure):
    _fields_ = [("means", _vp), ("scales", _vp), ("quats", _vp), ("densities", _vp), ("features", _vp),
                ("num_prims", _i64), ("feature_size", _i32), ("color_mode", _i32), ("sh_degree", _i32),
                ("activation", _i32), ("sh_origin", ctypes.c_float * 3), ("num_views", _i32),
                ("view_origins", _vp), ("view_colors", _vp)]


class Camera(ctypes.Structure):
    _fields_ = [("width", _i32), ("height", _i32), ("model", _i32), ("distorted", _i32), ("fx", ctypes.c_float),
                ("fy", ctypes.c_float), ("cx", ctypes.c_float), ("cy", ctypes.c_float), ("dist", ctypes.c_float * 6),
                ("c2w", _vp)]


class PlyLayout(ctypes.Structure):
    _fields_ = [("num_columns", _i32), ("mean", _i32 * 3), ("scale", _i32 * 3), ("rot", _i32 * 4), ("f_dc", _i32 * 3),
                ("f_rest0", _i32), ("num_rest", _i32), ("opacity", _i32), ("density", _i32), ("activation", _i32)]


class Rays(ctypes.Structure):
    _fields_ = [("origins", _vp), ("directions", _vp), ("num_rays", _i64), ("tmin", ctypes.c_float),
                ("tmax", ctypes.c_float), ("max_iters", _i32), ("kbuf", _i32), ("traversal", _i32),
                ("proposal", _vp), ("view_index", _vp), ("tile", _i32), ("event_layout", _i32)]


class ForwardOut(ctypes.Structure):
    _fields_ = [("color", _vp), ("states", _vp), ("iters", _vp), ("tri_collection", _vp), ("touch_count", _vp),
                ("initial_drgb", _vp), ("ray_state64", _vp), ("ray_order", _vp), ("workspace", _vp),
                ("workspace_bytes", ctypes.c_size_t), ("last_dirac", _vp)]


class BackwardIO(ctypes.Structure):
    _fields_ = [("grad_out", _vp), ("iters", _vp), ("tri_collection", _vp), ("states", _vp), ("ray_state64", _vp),
                ("ray_order", _vp), ("wcts", _vp), ("num_wcts", _i64), ("dmeans", _vp), ("dscales", _vp), ("dquats", _vp),
                ("ddensities", _vp), ("dfeatures", _vp), ("drayo", _vp), ("drayd", _vp), ("dmeans2D", _vp),
                ("dinitial_drgb", _vp), ("workspace", _vp), ("workspace_bytes", ctypes.c_size_t),
                ("view_color_grads", _vp)]


# Every function the header declares, with its ctypes signature.
SIGNATURES = {
    "ever_profile_enable": ([ctypes.c_int], ctypes.c_int),
    "ever_profile_read": ([ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "ever_mlp_param_count": ([], _i64),
    "ever_mlp_workspace_bytes": ([_i64], ctypes.c_size_t),
    "ever_mlp_sync_bf16": ([_vp, _vp, _vp], ctypes.c_int),
    "ever_mlp_infer": ([_vp, _vp, _vp, _vp, _i64, ctypes.c_float, ctypes.c_float, _vp, _vp, _vp, ctypes.c_size_t, _vp],
                       ctypes.c_int),
    "ever_mlp_train_step": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                             _i32, _vp, ctypes.c_size_t, _vp, _vp, _vp], ctypes.c_int),
    "ever_gemm_bf16": ([_vp, _i64, _i64, _vp, _i64, _i64, _i64, _i64, _i64, _vp, _i32, _vp, _i32, _vp, _vp],
                       ctypes.c_int),
    "ever_abi_version": ([], ctypes.c_int),
    "ever_struct_sizes": ([ctypes.POINTER(_i64)], ctypes.c_int),
    "ever_last_error": ([], ctypes.c_char_p),
    "ever_scene_layout": ([_i64, ctypes.POINTER(_i64)], ctypes.c_int),
    "ever_scene_workspace_bytes": ([_i64], ctypes.c_size_t),
    "ever_build_scene": ([ctypes.POINTER(Prims), _vp, ctypes.c_size_t, _vp], ctypes.c_int),
    "ever_update_colors": ([ctypes.POINTER(Prims), _vp, ctypes.c_size_t, _vp], ctypes.c_int),
    "ever_forward_workspace_bytes": ([_i64], ctypes.c_size_t),
    "ever_forward_workspace_bytes_for": ([_i64, _i32], ctypes.c_size_t),
    "ever_trace_forward": ([ctypes.POINTER(Prims), _vp, ctypes.c_size_t, ctypes.POINTER(Rays),
                            ctypes.POINTER(ForwardOut), _vp], ctypes.c_int),
    "ever_backward_workspace_bytes": ([_i64, _i64, _i64], ctypes.c_size_t),
    "ever_trace_backward": ([ctypes.POINTER(Prims), _vp, ctypes.c_size_t, ctypes.POINTER(Rays),
                             ctypes.POINTER(BackwardIO), _vp], ctypes.c_int),
    "ever_backward_recompute_workspace_bytes": ([_i64, _i64, _i64, _i32, _i32, _i64], ctypes.c_size_t),
    "ever_trace_backward_recompute": ([ctypes.POINTER(Prims), _vp, ctypes.c_size_t, ctypes.POINTER(Rays),
                                       ctypes.POINTER(BackwardIO), _i64, _vp], ctypes.c_int),
    "ever_scene_from_records": ([_vp, _i64, ctypes.POINTER(PlyLayout), _vp, _vp, _vp, _vp, _vp, _vp, _vp],
                                ctypes.c_int),
    "ever_morton_workspace_bytes": ([_i64], ctypes.c_size_t),
    "ever_morton_order": ([_vp, _i64, _vp, _vp, ctypes.c_size_t, _vp], ctypes.c_int),
    "ever_morton_order_noalloc": ([_vp, _i64, _vp, _vp], ctypes.c_int),
    "ever_camera_rays": ([ctypes.POINTER(Camera), _vp, _vp, _vp, _vp], ctypes.c_int),
    "ever_eval_sh": ([_vp, _vp, _vp, _i64, _i32, _i32, _vp, _vp], ctypes.c_int),
    "ever_eval_sh_backward": ([_vp, _vp, _vp, _vp, _i64, _i32, _i32, _vp, _vp], ctypes.c_int),
}

_lib = None


class EverError(RuntimeError):
    pass


def lib():
    """Load libever_amd.so; raises if it is absent (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EverError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        sizes = (_i64 * 4)()
        L.ever_struct_sizes(sizes)
        mine = [ctypes.sizeof(Prims), ctypes.sizeof(Rays), ctypes.sizeof(ForwardOut), ctypes.sizeof(BackwardIO)]
        if list(sizes) != mine:
            raise EverError(f"ABI struct size mismatch: library {list(sizes)} vs binding {mine}")
        _lib = L
    return _lib


def check(rc, what):
    if rc != EVER_OK:
        msg = lib().ever_last_error().decode()
        raise EverError(f"{what} failed ({rc}): {msg}")


def ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_ptr(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


LAYOUT_NAMES = ("header", "aabb", "fprim", "keys0", "vals0", "keys1", "vals1", "hist", "nodes", "nodes4", "range",
                "sbox", "pre", "suf", "tab", "rec", "colr", "cand", "bpart", "total")


def scene_layout(n):
    """Byte offsets of the scene workspace regions (ever_scene_layout)."""
    o = (_i64 * len(LAYOUT_NAMES))()
    check(lib().ever_scene_layout(n, o), "ever_scene_layout")
    return dict(zip(LAYOUT_NAMES, list(o)))
