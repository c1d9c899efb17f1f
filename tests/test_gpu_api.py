"""The drop-in API (splinetracers/fast_ellipsoid_splinetracer.py:37-256 of the reference): signature,
outputs, extras, autograd wiring, input validation, and that the HIP library is what runs."""
import numpy as np
import pytest
import torch

from ever_amd import scenes
from ever_amd.splinetracers import fast_ellipsoid_splinetracer as fes
from gpu_util import oracle_forward

pytestmark = pytest.mark.gpu
D = torch.device("cuda:0")


def inputs(n=500, w=32):
    sc = scenes.uniform_scene(n, 0, seed=1)
    o, d = scenes.pinhole_rays(w, w, (0.0, 0.0, -3.0))
    return sc, o.to(D), d.to(D)


def test_trace_rays_outputs_and_extras():
    sc, o, d = inputs()
    ps = [sc[k].to(D) for k in ("means", "scales", "quats", "densities", "features")]
    ps[3] = ps[3].reshape(-1, 1)  # (N,1) densities as the reference's callers pass
    out, ex = fes.trace_rays(*ps, o, d, 0.0, 100.0, 3, None, None, 500, True)
    assert out.shape == (o.shape[0], 5) and out.dtype == torch.float32
    assert set(ex) == {"tri_collection", "iters", "opacity", "touch_count", "distortion_loss", "saved"}
    assert ex["tri_collection"].shape == (o.shape[0] * 500,)
    it = ex["iters"].cpu().numpy()
    tri = ex["tri_collection"].view(500, -1).cpu().numpy()
    tail = np.arange(500)[:, None] >= it[None, :]
    assert np.all(tri[tail] == 0)  # zero-filled beyond each ray's events like the reference's torch::zeros
    assert torch.equal(ex["opacity"], out[:, 3]) and torch.equal(ex["distortion_loss"], out[:, 4])
    s = ex["saved"]
    assert s.states.shape == (o.shape[0], 16) and s.iters.shape == (o.shape[0],)
    assert int(ex["touch_count"].sum()) == int(ex["iters"].sum())
    ref = oracle_forward(sc, o.cpu(), d.cpu())
    np.testing.assert_allclose(out[:, :4].cpu().numpy(), ref["color"][:, :4], atol=1e-4)
    assert fes.trace_rays.uses_density


def test_autograd_gradients_shapes_and_none():
    sc, o, d = inputs()
    ps = [sc[k].to(D).requires_grad_(True) for k in ("means", "scales", "quats", "densities", "features")]
    m2 = torch.zeros((ps[0].shape[0], 2), device=D, requires_grad=True)
    out = fes.trace_rays(*ps, o, d, 0.0, 100.0, 3, m2, None)
    out.sum().backward()
    for p in ps:
        assert p.grad is not None and p.grad.shape == p.shape and torch.isfinite(p.grad).all()
    assert m2.grad is None  # no wcts: no means2D gradient (fast_ellipsoid_splinetracer.py:203)


def test_float32_finite_differences_like_reference_gradcheck():
    # tests/grad_check.py:64 style: FD in fp32 with eps 1e-3 against the analytic gradient, loose tolerance
    sc = scenes.uniform_scene(40, 0, seed=3)
    sc["means"][:, 2] += 0.5
    o, d = scenes.pinhole_rays(6, 6, (0.0, 0.0, -3.0))
    o, d = o.to(D), d.to(D)
    ps = [sc[k].to(D).requires_grad_(True) for k in ("means", "scales", "quats", "densities", "features")]
    g = scenes.upstream_grad(o.shape[0], seed=2).to(D)
    (fes.trace_rays(*ps, o, d, 0.0, 100.0) * g).sum().backward()
    rng = np.random.default_rng(0)
    for pi in range(5):
        p = ps[pi]
        for _ in range(6):
            idx = tuple(int(rng.integers(s)) for s in p.shape)
            base = p.detach().clone()
            vals = []
            for sgn in (1, -1):
                x = base.clone()
                x[idx] += sgn * 1e-3
                args = [q.detach() if j != pi else x for j, q in enumerate(ps)]
                vals.append(float((fes.trace_rays(*args, o, d, 0.0, 100.0) * g).sum()))
            fd = (vals[0] - vals[1]) / 2e-3
            an = float(p.grad[idx])
            assert abs(fd - an) <= 1e-2 + 2e-2 * abs(an), (pi, idx, fd, an)


def test_input_validation_errors():
    sc, o, d = inputs()
    ps = [sc[k].to(D) for k in ("means", "scales", "quats", "densities", "features")]
    with pytest.raises(RuntimeError, match="last dimension"):
        fes.trace_rays(ps[0], ps[1], ps[2][:, :3].contiguous(), *ps[3:], o, d)
    with pytest.raises(RuntimeError, match="same 0 dimension"):
        fes.trace_rays(ps[0], ps[1][:-1], *ps[2:], o, d)
    with pytest.raises(RuntimeError, match="float32"):
        fes.trace_rays(ps[0].double(), *ps[1:], o, d)
    with pytest.raises(RuntimeError, match="contiguous"):
        fes.trace_rays(*ps, o.t().contiguous().t(), d)


def test_native_library_is_loaded():
    import os
    sc, o, d = inputs()
    fes.trace_rays(*[sc[k].to(D) for k in ("means", "scales", "quats", "densities", "features")], o, d)
    maps = open("/proc/self/maps").read()
    assert "libever_amd.so" in maps
    assert os.path.basename(fes.C.LIB_PATH) == "libever_amd.so"
