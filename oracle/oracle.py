"""ctypes wrapper of the CPU oracle (ever_oracle.cpp) — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, as the checker or as
the timed CPU baseline. The product package (ever_amd/) never imports it.

Functions follow the reference files cited in ever_oracle.cpp; the argument meaning mirrors
splinetracers/fast_ellipsoid_splinetracer.py:222-254 (trace_rays) and :105-219 (backward).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libever_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.c_int64


def build():
    """Compile the oracle with the committed Makefile (g++, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for name, fp in (("oracle_forward_f64", _dp), ("oracle_forward_f32", _fp)):
            f = getattr(L, name)
            sc = ctypes.c_double if fp is _dp else ctypes.c_float
            args = [fp] * 5 + [_i64, ctypes.c_int, fp, fp, _i64, sc, sc, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, fp, fp, _ip, _ip, _ip, fp, _ip, _ip]
            if fp is _fp:
                args.append(_i64)
            f.argtypes = args
            f.restype = ctypes.c_int
        for name, fp in (("oracle_backward_f64", _dp), ("oracle_backward_f32", _fp)):
            f = getattr(L, name)
            sc = ctypes.c_double if fp is _dp else ctypes.c_float
            f.argtypes = [fp] * 5 + [_i64, ctypes.c_int, fp, fp, _i64, sc, ctypes.c_int, fp, _ip, _ip, fp, _ip,
                                     ctypes.c_int32, fp, _i64] + [fp] * 9 + [ctypes.c_int]
            f.restype = ctypes.c_int
        L.oracle_eval_sh_f64.argtypes = [_dp, _dp, _i64, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.oracle_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _arr(x, dtype):
    a = np.ascontiguousarray(np.asarray(x, dtype=dtype))
    return a


def _ptr(a):
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.float32:
        return a.ctypes.data_as(_fp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


def num_threads():
    return lib().oracle_num_threads()


def forward(means, scales, quats, densities, features, rayo, rayd, tmin=0.0, tmax=1000.0, max_iters=500,
            kbuf=16, use_bvh=False, safe=False, dtype=np.float64, ray_stride=1, with_tri=True):
    """Reference forward (shaders.slang:171-235). Returns a dict mirroring trace_rays' outputs and extras."""
    dt = np.dtype(dtype)
    means, scales, quats = _arr(means, dt), _arr(scales, dt), _arr(quats, dt)
    densities = _arr(densities, dt).reshape(-1)
    features = _arr(features, dt)
    n = means.shape[0]
    d = features.shape[1] if features.ndim == 3 else 1
    rayo, rayd = _arr(rayo, dt), _arr(rayd, dt)
    nr = rayo.shape[0]
    color = np.zeros((nr, 5), dt)
    states = np.zeros((nr, 16), dt)
    iters = np.zeros(nr, np.int32)
    tri = np.zeros(nr * max_iters, np.int32) if with_tri else None
    touch = np.zeros(n, np.int32)
    init = np.zeros((nr, 4), dt)
    inds = np.zeros(max(n, 1), np.int32)
    cnt = np.zeros(1, np.int32)
    L = lib()
    args = [_ptr(means), _ptr(scales), _ptr(quats), _ptr(densities), _ptr(features), n, d, _ptr(rayo), _ptr(rayd),
            nr, tmin, tmax, max_iters, kbuf, int(use_bvh), int(safe), _ptr(color), _ptr(states), _ptr(iters),
            _ptr(tri), _ptr(touch), _ptr(init), _ptr(inds), _ptr(cnt)]
    if dt == np.float64:
        assert ray_stride == 1
        L.oracle_forward_f64(*args)
    else:
        L.oracle_forward_f32(*args, ray_stride)
    return dict(color=color, states=states, iters=iters, tri_collection=tri, touch_count=touch,
                initial_drgb=init, initial_inds=inds[:cnt[0]].copy())


def backward(means, scales, quats, densities, features, rayo, rayd, tmin, max_iters, grad_out, fwd, wcts=None,
             clamp=True, dtype=np.float64):
    """Reference backward (backwards_kernel.slang:259-428 + clamps of fast_ellipsoid_splinetracer.py:201-219)."""
    dt = np.dtype(dtype)
    means, scales, quats = _arr(means, dt), _arr(scales, dt), _arr(quats, dt)
    densities = _arr(densities, dt).reshape(-1)
    features = _arr(features, dt)
    n = means.shape[0]
    d = features.shape[1] if features.ndim == 3 else 1
    rayo, rayd = _arr(rayo, dt), _arr(rayd, dt)
    nr = rayo.shape[0]
    g = _arr(grad_out, dt)
    iters = _arr(fwd["iters"], np.int32)
    tri = _arr(fwd["tri_collection"], np.int32)
    init = _arr(fwd["initial_drgb"], dt)
    inds = _arr(fwd["initial_inds"], np.int32)
    if inds.size == 0:
        inds = np.zeros(1, np.int32)
        cnt = 0
    else:
        cnt = int(fwd["initial_inds"].size)
    wct = None if wcts is None else _arr(wcts, dt).reshape(-1, 16)
    out = dict(dmeans=np.zeros((n, 3), dt), dscales=np.zeros((n, 3), dt), dquats=np.zeros((n, 4), dt),
               ddensities=np.zeros(n, dt), dfeatures=np.zeros(features.shape, dt), drayo=np.zeros((nr, 3), dt),
               drayd=np.zeros((nr, 3), dt), dmeans2D=np.zeros((n, 2), dt), dinitial_drgb=np.zeros((nr, 4), dt))
    f = lib().oracle_backward_f64 if dt == np.float64 else lib().oracle_backward_f32
    f(_ptr(means), _ptr(scales), _ptr(quats), _ptr(densities), _ptr(features), n, d, _ptr(rayo), _ptr(rayd), nr,
      tmin, max_iters, _ptr(g), _ptr(iters), _ptr(tri), _ptr(init), _ptr(inds), cnt, _ptr(wct),
      0 if wct is None else wct.shape[0], _ptr(out["dmeans"]), _ptr(out["dscales"]), _ptr(out["dquats"]),
      _ptr(out["ddensities"]), _ptr(out["dfeatures"]), _ptr(out["drayo"]), _ptr(out["drayd"]),
      _ptr(out["dmeans2D"]) if wct is not None else None, _ptr(out["dinitial_drgb"]), int(clamp))
    if wct is None:
        out["dmeans2D"] = None
    return out


def eval_sh(means, features, deg, origin):
    """Per-primitive SH colour at normalize(mean - origin), unclamped (sh_kernel.slang:19-42), FP64."""
    means = _arr(means, np.float64)
    features = _arr(features, np.float64)
    origin = _arr(origin, np.float64).reshape(3)
    n = means.shape[0]
    out = np.zeros((n, 3), np.float64)
    lib().oracle_eval_sh_f64(_ptr(means), _ptr(features), n, deg, features.shape[1], _ptr(origin), _ptr(out))
    return out
