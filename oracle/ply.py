"""Restatement of the reference viewer's PLY scene preparation — TEST INFRASTRUCTURE ONLY.

Only tests/ import this module, as the checker of ever_amd.ply (ever_scene_from_records, ever_morton_order).
Follows splinetracers/fast_ellipsoid_splinetracer/ply_file_loader.cc:150-193 (sigmoid, inv_opacity, softplus,
NormalizeSceneData) with the loader's float32 roundings, and :72-148 (SortComponentsToMortonOrder), whose float32
arithmetic is reproduced operation for operation so that codes are exact. Pinned in tests/test_ply_oracle.py by
closed forms of the activations, by de-interleaving the codes back to the quantised coordinates, and by brute-force
ordering.
"""
import numpy as np

FLT_MIN = np.float32(np.finfo(np.float32).tiny)
FLT_MAX = np.float32(np.finfo(np.float32).max)


def _f(x):
    return np.asarray(x, np.float64).astype(np.float32).astype(np.float64)


def activate_3dgs(scales_raw, opacity, quats_raw):
    """(scales, densities, quats) of NormalizeSceneData for raw 3DGS values. Every step the loader evaluates in
    float (its sigmoid and softplus, 1 - alpha, the final division) is rounded to float32 here too, exp and log
    themselves in FP64: the loader's softplus log(1 + exp(x)) loses most of its digits to the float sum for
    negative x, and that loss is part of the result."""
    s = _f(scales_raw)
    sp = np.where(s > 5, s, _f(np.log(_f(1.0 + _f(np.exp(np.minimum(s, 5.0)))))))
    scales = _f(np.minimum(sp + 3e-4, 25.0))
    norm = np.minimum(np.min(_f(2 * scales), axis=1), 9999.0)
    sig = _f(1.0 / _f(1.0 + _f(np.exp(-_f(opacity)))))
    a = _f(0.99 * sig)
    dens = _f(_f(np.maximum(0.0, -np.log(np.maximum(_f(1.0 - a), 1e-10)))) / norm)
    q = _f(quats_raw)
    q = q / (q * q).sum(1, keepdims=True)  # divided by the squared length, as the loader does
    return scales, dens, q


def morton_codes(means):
    """63-bit codes (uint64) with the loader's float32 arithmetic: bounds from FLT_MAX / FLT_MIN starts,
    rel = (m - lo) / (hi - lo), int(float(2^21 - 1) * rel), bit b of x, y, z at 3b, 3b+1, 3b+2."""
    m = np.asarray(means, np.float32)
    lo = np.minimum(m.min(0), FLT_MAX) if len(m) else np.full(3, FLT_MAX, np.float32)
    hi = np.maximum(m.max(0), FLT_MIN) if len(m) else np.full(3, FLT_MIN, np.float32)
    code = np.zeros(len(m), np.uint64)
    with np.errstate(invalid="ignore", divide="ignore"):
        rel = (m - lo) / (hi - lo)
        sc = np.float32(2097151.0) * rel
    ok = np.isfinite(sc)
    u = np.where(ok, sc, 0).astype(np.int64).astype(np.uint64)
    for a in range(3):
        for b in range(21):
            code |= ((u[:, a] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + a)
    return code


def morton_order(means):
    """Source index of each row sorted by code, ties in index order."""
    return np.argsort(morton_codes(means), kind="stable").astype(np.int32)
