"""The reference's compiled-module API (`sp`, py_binding.cpp:261-282) as its inference caller uses it
(integration_utils/gaussian_renderer/fast_renderer.py:50-63, :112-119): a trace through sp equals the autograd
path's trace of the same inputs, new features through set_features + update_model equal a fresh build with
those features, enable_backward=False gives the same image with the event list and touch counts left at zero,
and the start-inside candidate list equals the oracle's prefilter set."""
import numpy as np
import pytest
import torch

from ever_amd import scenes
from ever_amd.splinetracers import fast_ellipsoid_splinetracer as fes
from ever_amd.splinetracers.fast_ellipsoid_splinetracer import sp
from gpu_util import KEYS, oracle_forward

pytestmark = pytest.mark.gpu
D = torch.device("cuda:0")
EYE = (0.0, 0.0, -4.0)


def _tensors(sc):
    return [sc[k].to(D).contiguous() for k in KEYS]


def _sp_setup(t, enable_backward):
    means, scales, quats, dens, feats = t
    ctx = sp.OptixContext(D)
    prims = sp.Primitives(D)
    half = torch.cat([means, scales, quats], dim=1).half().contiguous()
    prims.add_primitives(means, scales, quats, half, dens, feats)
    gas = sp.GAS(ctx, D, prims, True, False, True)
    fwd = sp.Forward(ctx, D, prims, enable_backward)
    return prims, gas, fwd


@pytest.mark.parametrize("enable_backward", [True, False])
def test_sp_trace_equals_trace_rays(enable_backward):
    print(f"case: sp vs trace_rays, clustered 20k, 160x120, enable_backward {enable_backward}", flush=True)
    sc = scenes.clustered_scene(20_000, 0, seed=21)
    t = _tensors(sc)
    o, d = scenes.pinhole_rays(160, 120, EYE)
    o, d = o.to(D).contiguous(), d.to(D).contiguous()
    ref, ex = fes.trace_rays(*t, o, d, 0.0, 100.0, max_iters=200, return_extras=True)
    _, gas, fwd = _sp_setup(t, enable_backward)
    out = fwd.trace_rays(gas, o, d, 0.0, 100.0, 200, 1000)
    nr = o.shape[0]
    assert set(out) == {"color", "saved", "tri_collection", "initial_drgb", "initial_touch_inds",
                        "initial_touch_count"}
    assert out["color"].shape == (nr, 4) and out["color"].is_contiguous()
    assert torch.equal(out["color"], ref[:, :4])
    s = out["saved"]
    assert torch.equal(s.iters, ex["iters"])
    assert torch.equal(s.states, ex["saved"].states)
    assert s.diracs.shape == (nr, 4) and s.faces.shape == (nr,)
    assert out["tri_collection"].shape == (nr * 200,) and out["tri_collection"].dtype == torch.int32
    if enable_backward:
        assert torch.equal(out["tri_collection"], ex["tri_collection"])
        assert torch.equal(s.touch_count, ex["touch_count"])
    else:
        assert int(out["tri_collection"].abs().max()) == 0 and int(s.touch_count.abs().max()) == 0
    # rays start outside every primitive's bounds: no candidates, zero start-inside colour
    assert int(out["initial_touch_count"][0]) == 0 and int(out["initial_touch_inds"].abs().max()) == 0
    assert float(out["initial_drgb"].abs().max()) == 0.0


def test_update_model_equals_fresh_build():
    print("case: set_features + update_model vs fresh GAS", flush=True)
    sc = scenes.clustered_scene(20_000, 0, seed=22)
    t = _tensors(sc)
    o, d = scenes.pinhole_rays(128, 96, EYE)
    o, d = o.to(D).contiguous(), d.to(D).contiguous()
    prims, gas, fwd = _sp_setup(t, False)
    a0 = fwd.trace_rays(gas, o, d, 0.0, 100.0, 200, 1000)["color"].clone()
    g = torch.Generator().manual_seed(3)
    new = (torch.randn(t[4].shape, generator=g) * 0.8).to(D).contiguous()
    prims.set_features(new)
    # the Forward keeps its features until update_model
    b0 = fwd.trace_rays(gas, o, d, 0.0, 100.0, 200, 1000)["color"]
    assert torch.equal(a0, b0)
    fwd.update_model(prims)
    b = fwd.trace_rays(gas, o, d, 0.0, 100.0, 200, 1000)["color"]
    _, gas2, fwd2 = _sp_setup(t[:4] + [new], False)
    c = fwd2.trace_rays(gas2, o, d, 0.0, 100.0, 200, 1000)["color"]
    assert torch.equal(b, c) and not torch.equal(b, a0)
    # a second Forward with the old features on the same GAS still traces the old colours
    fwd_old = sp.Forward(sp.OptixContext(D), D, _sp_setup(t, False)[0], False)
    assert torch.equal(fwd_old.trace_rays(gas, o, d, 0.0, 100.0, 200, 1000)["color"], a0)
    # the density channel is untouched by a colour update
    assert torch.equal(b[:, 3], a0[:, 3])


def test_start_inside_candidates_match_oracle():
    print("case: sp start-inside candidates, 300 prims, tmin 0.01", flush=True)
    sc = scenes.uniform_scene(300, 0, seed=4)
    rng = np.random.default_rng(4)
    o = torch.tensor(np.tile(sc["means"][0].numpy(), (256, 1)) + rng.normal(0, 1e-3, (256, 3)), dtype=torch.float32)
    d = torch.tensor(rng.normal(size=(256, 3)), dtype=torch.float32)
    d = d / d.norm(dim=1, keepdim=True)
    t = _tensors(sc)
    _, gas, fwd = _sp_setup(t, True)
    out = fwd.trace_rays(gas, o.to(D).contiguous(), d.to(D).contiguous(), 0.01, 100.0, 500, 1000)
    ref = oracle_forward(sc, o, d, tmin=0.01)
    n = int(out["initial_touch_count"][0])
    inds = out["initial_touch_inds"].cpu().numpy()
    assert n == len(ref["initial_inds"]) > 0
    assert sorted(inds[:n].tolist()) == sorted(np.asarray(ref["initial_inds"]).tolist())
    assert (inds[n:] == 0).all()
    np.testing.assert_allclose(out["initial_drgb"].cpu().numpy(), ref["initial_drgb"], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(out["color"].cpu().numpy(), ref["color"][:, :4], atol=1e-4, rtol=1e-4)
