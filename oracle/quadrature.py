"""FP64 midpoint-quadrature renderer — TEST INFRASTRUCTURE ONLY (an independent cross-check of the oracle).

numpy restatement of the reference's ground-truth renderer used by its tests:
  query_ellipsoid          splinetracers/quad.py:54-61   (normalized quaternion, conjugate rotation, max(s,1e-8))
  sum + safe_div + clip    splinetracers/quad.py:93-97
  trace_rays               splinetracers/quad.py:67-115  (2^16 samples on [tmin, tmax], SH2RGB of the DC term)
  render_quadrature        jaxutil/quadrature.py:182-205 ; weights :169-180 ; distortion :148-160
  safe_div / safe_log      jaxutil/safe_math.py:130-156, generate_safe_fn (log on [tiny, max])
"""
import numpy as np

_TINY32 = float(np.finfo(np.float32).tiny)
_MAX32 = float(np.finfo(np.float32).max)
_MIN32 = float(np.finfo(np.float32).min)
SH_C0 = 0.28209479177387814


def quat_to_rot_T(q):
    """jquatToMat3 (utils/math_util.py:79-100): returns R(q)^T for q = (w, x, y, z)."""
    r, x, y, z = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - r * z), 2 * (x * z + r * y)],
                  [2 * (x * y + r * z), 1 - 2 * (x * x + z * z), 2 * (y * z - r * x)],
                  [2 * (x * z - r * y), 2 * (y * z + r * x), 1 - 2 * (x * x + y * y)]])
    return R.T


def _safe_div(n, d):
    r = np.clip(n / np.where(np.abs(d) < _TINY32, _TINY32, d), _MIN32, _MAX32)
    return np.where(np.abs(d) < _TINY32, 0.0, r)


def render(means, scales, quats, densities, features, rayo, rayd, tmin=0.0, tmax=3.0, num_quad=2 ** 16):
    """Returns (R, 5) = [RGB, alpha, distortion] per ray, FP64."""
    means = np.asarray(means, np.float64).reshape(-1, 3)
    scales = np.asarray(scales, np.float64).reshape(-1, 3)
    quats = np.asarray(quats, np.float64).reshape(-1, 4)
    dens = np.asarray(densities, np.float64).reshape(-1)
    cols = SH2RGB(np.asarray(features, np.float64).reshape(len(means), -1, 3)[:, 0, :])
    rayo = np.asarray(rayo, np.float64).reshape(-1, 3)
    rayd = np.asarray(rayd, np.float64).reshape(-1, 3)
    tdist = np.linspace(tmin, tmax, num_quad + 1)
    t_avg = 0.5 * (tdist[1:] + tdist[:-1])
    t_delta = np.diff(tdist)
    out = np.zeros((rayo.shape[0], 5))
    for r in range(rayo.shape[0]):
        xs = rayo[r][None, :] + t_avg[:, None] * rayd[r][None, :]
        tot_d = np.zeros(num_quad)
        tot_c = np.zeros((num_quad, 3))
        for i in range(means.shape[0]):
            qn = quats[i] / np.sqrt(np.sum(quats[i] ** 2))
            R = quat_to_rot_T(qn)
            sc = ((xs - means[i][None, :]) @ R.T) / np.maximum(scales[i][None, :], 1e-8)
            inside = np.linalg.norm(sc, axis=-1) < 1
            tot_d += np.where(inside, dens[i], 0.0)
            tot_c += np.where(inside[:, None], dens[i] * cols[i][None, :], 0.0)
        avg_c = np.clip(_safe_div(tot_c, tot_d[:, None]), 0, None)
        dd = tot_d * t_delta
        log_trans = -np.concatenate([[0.0], np.cumsum(dd[:-1])])
        log1mexp = np.log(np.clip(1 - np.exp(-dd), _TINY32, _MAX32))
        w = np.exp(log1mexp + log_trans)
        out[r, :3] = np.sum(w[:, None] * avg_c, axis=0)
        out[r, 3] = np.sum(w)
        ut = t_avg
        # distortion (jaxutil/quadrature.py:148-160) restated as an O(n) sweep of the O(n^2) double sum
        cw = np.cumsum(w)
        cwu = np.cumsum(w * ut)
        inter = 2 * np.sum(w[1:] * (ut[1:] * cw[:-1] - cwu[:-1]))
        intra = np.sum(w ** 2 * t_delta) / 3
        out[r, 4] = inter + intra
    return out


def SH2RGB(x):
    """quad.py:64-65"""
    return SH_C0 * x + 0.5


def RGB2SH(rgb):
    """third_party/sh_util.py:122-123"""
    return (rgb - 0.5) / SH_C0
