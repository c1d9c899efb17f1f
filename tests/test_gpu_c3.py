"""Clustered-scene differentiation checks: the adjoint with recomputed events (no event list kept between forward
and backward) equals the adjoint with stored events, and the analytic gradient of a scalar loss matches FP64
central finite differences of the oracle on 256 random parameters (means, scales, quaternions, densities, SH
features) of the primitives the rays touch."""
import numpy as np
import pytest
import torch

from ever_amd import scenes
from ever_amd import tracer as T
from ever_amd.splinetracers import fast_ellipsoid_splinetracer as fes
from oracle import oracle

pytestmark = pytest.mark.gpu
D = torch.device("cuda:0")
EYE = (0.0, 0.0, -4.0)


def _params(sc):
    return [sc[k].to(D).contiguous().requires_grad_(True) for k in ("means", "scales", "quats", "densities",
                                                                    "features")]


def _grads(params, o, d, g):
    out = fes.trace_rays_sh(*params, o, d, EYE, 3, 0.0, 100.0, max_iters=500)
    torch.autograd.backward(out, g)
    gr = [p.grad.detach().clone() for p in params]
    for p in params:
        p.grad = None
    return out.detach(), gr


@pytest.mark.parametrize("chunk", [10_000, 65_536, 262_144])
def test_recomputed_events_equal_stored_events(chunk):
    # 57,600 rays: six chunks with a ragged last one (per-ray walk), one chunk of 65,536 (per-wave walk, sorted),
    # one chunk larger than the call
    print(f"case: clustered 100k, 320x180, stored vs recomputed events, chunks of {chunk} rays", flush=True)
    sc = scenes.clustered_scene(100_000, 3, seed=9)
    o, d = scenes.pinhole_rays(320, 180, EYE)
    o, d = o.to(D).contiguous(), d.to(D).contiguous()
    g = scenes.upstream_grad(o.shape[0], seed=2).to(D)
    params = _params(sc)
    try:
        fes.set_event_mode("stored")
        a, ga = _grads(params, o, d, g)
        fes.set_event_mode("recompute", chunk_rays=chunk)
        b, gb = _grads(params, o, d, g)
    finally:
        fes.set_event_mode("stored", chunk_rays=T.RECOMPUTE_CHUNK)
    assert torch.equal(a, b)
    for x, y in zip(ga, gb):
        assert float((x - y).norm()) <= 1e-6 * float(x.norm()) + 1e-12


def _peak_step(params, o, d, g):
    """Peak device memory allocated by one forward + backward above what was allocated before it."""
    torch.cuda.synchronize()
    base = torch.cuda.memory_allocated()
    torch.cuda.reset_peak_memory_stats()
    _, gr = _grads(params, o, d, g)
    torch.cuda.synchronize()
    return torch.cuda.max_memory_allocated() - base, gr


def test_recompute_bounds_memory_at_full_size():
    """1M clustered SH3 primitives at 1920x1080 (2,073,600 rays, max_iters 500): the stored event list is 4.15 GB;
    the recomputing adjoint holds one chunk's list instead, with the same gradients."""
    print("case: C3 1M clustered 1920x1080, peak memory stored vs recompute", flush=True)
    sc, o, d = scenes.config3()
    o, d = o.to(D).contiguous(), d.to(D).contiguous()
    g = scenes.upstream_grad(o.shape[0], seed=5).to(D)
    params = _params(sc)
    try:
        fes.set_event_mode("stored")
        peak_s, ga = _peak_step(params, o, d, g)
        fes.set_event_mode("recompute", chunk_rays=T.RECOMPUTE_CHUNK)
        peak_r, gb = _peak_step(params, o, d, g)
    finally:
        fes.set_event_mode("stored")
    print(f"peak above baseline: stored {peak_s / 1e9:.3f} GB, recompute {peak_r / 1e9:.3f} GB, "
          f"saved {(peak_s - peak_r) / 1e9:.3f} GB", flush=True)
    assert peak_s - peak_r >= 3.5e9
    for x, y in zip(ga, gb):
        assert float((x - y).norm()) <= 1e-6 * float(x.norm()) + 1e-12


def _colour_feature(means, feats):
    """The SH prepass of the oracle: eval_sh at the eye, softplus(10), RGB2SH -> (N,1,3) degree-0 features."""
    rgb = oracle.eval_sh(means, feats, 3, np.array(EYE))
    return ((np.logaddexp(0.0, 10 * rgb) / 10 - 0.5) / 0.28209479177387814).reshape(-1, 1, 3)


def _oracle_loss(sc, f0, o, d, g):
    """sum_r g_r . colour_r of the FP64 oracle for given degree-0 colour features. The safe-root form of the
    forward (rotated direction l2-normalised, tri-intersect.slang:332-344) is the function the reference's
    backward differentiates; the two agree in value and differ in the quaternion's radial derivative."""
    f = oracle.forward(sc["means"], sc["scales"], sc["quats"], sc["densities"], f0, o, d, 0.0, 100.0,
                       max_iters=500, use_bvh=True, safe=True)
    return float((f["color"] * g).sum())


def test_gradients_match_finite_differences_on_256_parameters():
    print("case: clustered 20k, 64x48, FD on 256 parameters", flush=True)
    sc = scenes.clustered_scene(20_000, 3, seed=10)
    o, d = scenes.pinhole_rays(64, 48, EYE)
    g = scenes.upstream_grad(o.shape[0], seed=3)
    params = _params(sc)
    _, grads = _grads(params, o.to(D).contiguous(), d.to(D).contiguous(), g.to(D))
    grads = [x.cpu().double().numpy() for x in grads]
    npsc = {k: sc[k].double().numpy().copy() for k in ("means", "scales", "quats", "densities", "features")}
    names = ["means", "scales", "quats", "densities", "features"]
    o64, d64, g64 = o.double().numpy(), d.double().numpy(), g.double().numpy()
    # parameters of primitives with a non-negligible gradient (the ones the rays touch)
    touched = np.nonzero(np.abs(grads[3]) > 1e-6)[0]
    assert len(touched) > 100
    rng = np.random.default_rng(4)
    picks = []
    while len(picks) < 256:
        k = int(rng.integers(0, 5))
        p = int(rng.choice(touched))
        flat = grads[k][p].reshape(-1)
        j = int(rng.integers(0, flat.size))
        if k == 4 and j >= 48:
            continue
        picks.append((k, p, j))
    # The colour gradient reaches the SH features only: as the reference's eval_sh backward (eval_sh.py:68), no
    # gradient flows to the means through the SH view direction. So the colours are held fixed while geometry is
    # perturbed and recomputed (at the unperturbed means) while features are.
    f0_fixed = _colour_feature(npsc["means"], npsc["features"])

    def loss():
        return _oracle_loss(npsc, f0_fixed, o64, d64, g64)

    def central(k, p, j, h):
        arr = npsc[names[k]]
        view = arr.reshape(arr.shape[0], -1)[p]  # a view into arr (densities are 1-D)
        x0 = view[j]
        vals = []
        for s in (1.0, -1.0):
            view[j] = x0 + s * h
            if k == 4:
                vals.append(_oracle_loss(npsc, _colour_feature(npsc["means"], npsc["features"]), o64, d64, g64))
            else:
                vals.append(loss())
        view[j] = x0
        return (vals[0] - vals[1]) / (2 * h)

    fd, fd2, an = [], [], []
    for k, p, j in picks:
        x0 = npsc[names[k]].reshape(npsc[names[k]].shape[0], -1)[p][j]
        h = 1e-5 * max(1.0, abs(x0))
        fd.append(central(k, p, j, h))
        fd2.append(central(k, p, j, 2 * h))
        an.append(grads[k][p].reshape(-1)[j])
    fd, fd2, an = np.array(fd), np.array(fd2), np.array(an)
    scale = np.median(np.abs(fd)) + 1e-12
    # central differences err by c*h^2: fd2 - fd = 3*c*h^2 measures it and Richardson's (4*fd - fd2)/3 removes it.
    # A parameter on a kink (grazing silhouette, two events changing order) has no derivative there, and one next
    # to a kink has a curvature the steps cannot resolve: the two step sizes disagree by more than 1e-3, and such
    # parameters are reported and left out
    smooth = np.abs(fd - fd2) <= 1e-3 * np.maximum(np.abs(fd), scale)
    smooth &= np.abs(an) < 999.0  # gradients clamped at +-1e3 (fast_ellipsoid_splinetracer.py:201-211)
    kinds = np.array([k for k, _, _ in picks])
    fd = (4.0 * fd - fd2) / 3.0
    err = np.abs(fd - an) / np.maximum(np.abs(fd), scale)
    print("FD rel err (smooth params): median", np.median(err[smooth]), "p95", np.percentile(err[smooth], 95),
          "max", err[smooth].max(), "; left out", int((~smooth).sum()), "of", len(fd), flush=True)
    for k in range(5):
        sel = smooth & (kinds == k)
        if sel.any():
            print("  ", names[k], int(sel.sum()), "median", np.median(err[sel]), "max", err[sel].max(), flush=True)
    assert smooth.sum() >= 240
    assert np.median(err[smooth]) < 1e-4 and err[smooth].max() <= 1e-3
