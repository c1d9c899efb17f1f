"""Pins the FP64 oracle (CPU): closed forms, the restated quadrature renderer, finite differences, invariants.

Scenario sources in the reference: tests/compare_quadrature.py:30-102 (quadrature scenes, atol=rtol=1e-4),
tests/grad_check.py:31-99 (gradient scenes), splinetracers/fast_ellipsoid_splinetracer/slang/shaders.slang:15-18
(termination constants).
"""
import math

import numpy as np
import pytest

from oracle import oracle
from oracle import quadrature as Q

SH_C0 = 0.28209479177387814


def one_ray():
    return np.array([[0.0, 0.0, 0.0]]), np.array([[0.0, 0.0, 1.0]])


def sphere(center, r, sigma, f=(0.4, -0.2, 0.1), quat=(1.0, 0.0, 0.0, 0.0)):
    return (np.array([center], float), np.full((1, 3), r), np.array([quat], float), np.array([sigma]),
            np.array([[f]], float))


def fwd(scene, rayo, rayd, tmin=0.0, tmax=100.0, **kw):
    return oracle.forward(*scene, rayo, rayd, tmin, tmax, **kw)


def test_sphere_through_centre_closed_form():
    r, s = 0.3, 2.0
    scene = sphere((1e-7, 0.0, 1.0), r, s)
    out = fwd(scene, *one_ray())
    c = SH_C0 * np.array([0.4, -0.2, 0.1]) + 0.5
    op = 1 - math.exp(-2 * s * r)
    np.testing.assert_allclose(out["color"][0, 3], op, rtol=1e-12)
    np.testing.assert_allclose(out["color"][0, :3], c * op, rtol=1e-12)
    assert out["iters"][0] == 2
    assert list(out["tri_collection"][:2]) == [1, 0]  # entry id 2*0+1 then exit 2*0


def test_chord_closed_form_off_centre_rotated_ellipsoid():
    # ellipsoid with axes (a, b, c) rotated 90 deg about z: local x maps to world y. A ray along +z at
    # world offset (0, y0) crosses local coordinates (y0 along local x): chord = 2 c sqrt(1 - (y0/a)^2).
    a, b, c, y0, s = 0.5, 0.2, 0.3, 0.25, 1.5
    q = (math.cos(math.pi / 4), 0.0, 0.0, math.sin(math.pi / 4))
    scene = (np.array([[0.0, 0.0, 2.0]]), np.array([[a, b, c]]), np.array([q]), np.array([s]),
             np.array([[[0.0, 0.0, 0.0]]]))
    out = fwd(scene, np.array([[0.0, y0, 0.0]]), np.array([[0.0, 0.0, 1.0]]))
    chord = 2 * c * math.sqrt(1 - (y0 / a) ** 2)
    np.testing.assert_allclose(out["color"][0, 3], 1 - math.exp(-s * chord), rtol=1e-10)


def test_nested_spheres_piecewise_closed_form():
    # outer sphere radius 0.5 (sigma1, c1) containing inner radius 0.2 (sigma2, c2), same centre on the ray.
    s1, s2 = 1.0, 3.0
    f1, f2 = (0.2, 0.1, -0.3), (-0.4, 0.5, 0.0)
    scene = (np.array([[1e-7, 0, 1.0], [1e-7, 0, 1.0]]), np.array([[0.5] * 3, [0.2] * 3]),
             np.array([[1.0, 0, 0, 0]] * 2), np.array([s1, s2]), np.array([[f1], [f2]]))
    out = fwd(scene, *one_ray())
    c1 = SH_C0 * np.array(f1) + 0.5
    c2 = SH_C0 * np.array(f2) + 0.5
    cm = (s1 * c1 + s2 * c2) / (s1 + s2)
    seg = [(0.3, s1, c1), (0.4, s1 + s2, cm), (0.3, s1, c1)]
    T, C = 1.0, np.zeros(3)
    for L, sg, col in seg:
        w = T * (1 - math.exp(-sg * L))
        C += w * col
        T *= math.exp(-sg * L)
    np.testing.assert_allclose(out["color"][0, :3], C, rtol=1e-10)
    np.testing.assert_allclose(out["color"][0, 3], 1 - T, rtol=1e-10)


def test_distortion_equals_pair_sum():
    # distortion = sum_{i>j} 2 w_i w_j (m_i - m_j), m = ladder(1000 * segment midpoint)
    rng = np.random.default_rng(3)
    n = 8
    scene = (np.c_[rng.uniform(-0.1, 0.1, (n, 2)), rng.uniform(0.5, 2.5, n)], rng.uniform(0.1, 0.4, (n, 3)),
             np.tile([1.0, 0, 0, 0], (n, 1)), rng.uniform(0.2, 1.5, n), rng.normal(0, 0.5, (n, 1, 3)))
    out = fwd(scene, *one_ray())
    tri = out["tri_collection"][: out["iters"][0]]
    # rebuild the segment list from the event order and the oracle's own roots through closed form
    ts, ws = [], []
    t_prev, D, logT = 0.0, 0.0, 0.0
    for tid in tri:
        p, face = tid // 2, tid % 2
        m, s = scene[0][p], scene[1][p]
        # sphere-aligned ellipsoid roots along +z from the origin
        h = 1 - (m[0] / s[0]) ** 2 - (m[1] / s[1]) ** 2
        t0, t1 = m[2] - s[2] * math.sqrt(h), m[2] + s[2] * math.sqrt(h)
        t = t0 if face == 1 else t1
        area = max(D * (t - t_prev), 0)
        ws.append(math.exp(-logT) * (1 - math.exp(-area)))
        ts.append((t + t_prev) / 2)
        logT += area
        D += scene[3][p] if face == 1 else -scene[3][p]
        t_prev = t
    x = np.array(ts) * 1000
    mm = 11 * np.sign(x) * (1 - (1 + np.abs(x) / 1.1) ** -0.1)
    w = np.array(ws)
    ref = sum(2 * w[i] * w[j] * (mm[i] - mm[j]) for i in range(len(w)) for j in range(i))
    np.testing.assert_allclose(out["color"][0, 4], ref, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("N", [1, 5, 10, 20, 40])
@pytest.mark.parametrize("density_multi", [0.01, 0.1, 1])
def test_quadrature_ray_origin_outside(N, density_multi):
    # tests/compare_quadrature.py:30-65, seeded
    rng = np.random.default_rng(1000 * N + int(100 * density_multi))
    scale = 0.5 * rng.random((N, 3))
    mean = rng.random((N, 3))
    mean[:, 0] *= 0.5
    mean[:, 1] *= 0.5
    mean[:, 2] += 0.5
    q = 2 * rng.random((N, 4)) - 1
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    dens = density_multi * rng.random((N, 1))
    feats = rng.random((N, 1, 3))
    ro, rd = one_ray()
    a = Q.render(mean, scale, q, dens, feats, ro, rd, 0, 3)[0, :4]
    b = oracle.forward(mean, scale, q, dens, feats, ro, rd, 0, 100)["color"][0, :4]
    np.testing.assert_allclose(a, b, atol=1e-4, rtol=1e-4)


@pytest.mark.parametrize("N", [1, 5, 10, 20, 40])
@pytest.mark.parametrize("density_multi", [0.01, 0.1, 1])
def test_quadrature_ray_origin_inside(N, density_multi):
    # tests/compare_quadrature.py:67-102, seeded; exercises the start-inside density
    rng = np.random.default_rng(7 + 1000 * N + int(100 * density_multi))
    scale = 0.5 * rng.random((N, 3))
    mean = 1.2 * rng.random((N, 3)) - 0.2
    mean[:, 0] *= 0.5
    mean[:, 1] *= 0.5
    q = 2 * rng.random((N, 4)) - 1
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    dens = density_multi * rng.random((N, 1))
    feats = rng.random((N, 1, 3))
    tmin = rng.random() * 0.3
    ro, rd = one_ray()
    a = Q.render(mean, scale, q, dens, feats, ro, rd, tmin, 3)[0, :4]
    b = oracle.forward(mean, scale, q, dens, feats, ro, rd, tmin, 100)["color"][0, :4]
    np.testing.assert_allclose(a, b, atol=1e-4, rtol=1e-4)


def test_termination_log_cutoff_and_max_iters():
    # a row of dense spheres: the ray stops once log-transmittance reaches 5.54 (shaders.slang:17,195)
    n = 12
    scene = (np.c_[np.full((n, 2), 1e-7), 0.5 + 0.5 * np.arange(n)], np.full((n, 3), 0.2),
             np.tile([1.0, 0, 0, 0], (n, 1)), np.full(n, 5.0), np.zeros((n, 1, 3)))
    out = fwd(scene, *one_ray())
    # each sphere contributes 5*0.4 = 2 of optical depth: cut after the 3rd sphere's exit (6 >= 5.54)
    assert out["iters"][0] == 6
    np.testing.assert_allclose(out["states"][0, 12], 6.0, rtol=1e-12)
    out2 = fwd(scene, *one_ray(), max_iters=3)
    assert out2["iters"][0] == 3
    np.testing.assert_allclose(out2["states"][0, 12], 2.0, rtol=1e-12)


def test_tmax_truncation_no_tail_segment():
    # events beyond tmax are not reported and no density is integrated past the last event
    scene = (np.array([[1e-7, 0, 1.0], [1e-7, 0, 3.0]]), np.full((2, 3), 0.25), np.tile([1.0, 0, 0, 0], (2, 1)),
             np.array([1.0, 1.0]), np.zeros((2, 1, 3)))
    out = fwd(scene, *one_ray(), tmax=2.0)
    assert out["iters"][0] == 2
    np.testing.assert_allclose(out["color"][0, 3], 1 - math.exp(-0.5), rtol=1e-12)
    # tmax inside the first sphere: its entry is reported (t0 <= tmax) and its exit comes with it
    out = fwd(scene, *one_ray(), tmax=0.9)
    assert out["iters"][0] == 2


def test_start_inside_density():
    # camera inside a sphere: the initial density carries it, its exit event closes it
    r, s = 0.5, 2.0
    scene = sphere((1e-7, 0.0, 0.1), r, s)
    out = fwd(scene, *one_ray())
    np.testing.assert_allclose(out["initial_drgb"][0, 0], s)
    assert out["iters"][0] == 1
    np.testing.assert_allclose(out["color"][0, 3], 1 - math.exp(-s * 0.6), rtol=1e-9)


def test_kbuffer_size_independent():
    rng = np.random.default_rng(5)
    n = 60
    scene = (np.c_[rng.uniform(-0.3, 0.3, (n, 2)), rng.uniform(0.5, 4.0, n)], rng.uniform(0.05, 0.3, (n, 3)),
             rng.normal(size=(n, 4)), rng.uniform(0.05, 0.4, n), rng.normal(0, 0.5, (n, 1, 3)))
    scene[2][:] /= np.linalg.norm(scene[2], axis=1, keepdims=True)
    rays = np.c_[rng.uniform(-0.1, 0.1, (16, 2)), np.zeros(16)], np.c_[rng.uniform(-0.1, 0.1, (16, 2)), np.ones(16)]
    a = fwd(scene, *rays, kbuf=16)
    for k in (1, 3, 8, 32):
        b = fwd(scene, *rays, kbuf=k)
        np.testing.assert_allclose(a["color"], b["color"], rtol=1e-12, atol=1e-14)
        np.testing.assert_array_equal(a["iters"], b["iters"])


def test_tied_events_across_kbuffer_boundary_processed_once():
    # two coincident spheres give bit-identical event depths; with a 1-slot k-buffer every event is a chunk
    # boundary, and the (t, key) cursor must still process all four events
    scene = (np.array([[1e-7, 0, 1.0], [1e-7, 0, 1.0]]), np.full((2, 3), 0.3), np.tile([1.0, 0, 0, 0], (2, 1)),
             np.array([1.0, 2.0]), np.array([[[0.1, 0.2, 0.3]], [[-0.3, 0.0, 0.4]]]))
    a = fwd(scene, *one_ray(), kbuf=1)
    b = fwd(scene, *one_ray(), kbuf=16)
    assert a["iters"][0] == b["iters"][0] == 4
    np.testing.assert_allclose(a["color"], b["color"], rtol=0, atol=0)
    np.testing.assert_allclose(a["color"][0, 3], 1 - math.exp(-3.0 * 0.6), rtol=1e-12)


def test_bvh_matches_brute_force():
    rng = np.random.default_rng(6)
    n = 400
    means = rng.uniform(-1, 1, (n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    scene = (means, np.exp(rng.uniform(np.log(0.02), np.log(0.2), (n, 3))), q,
             np.exp(rng.uniform(np.log(0.1), np.log(50), n)), rng.normal(0, 0.5, (n, 1, 3)))
    rd = rng.normal(size=(64, 3)) * 0.2 + np.array([0, 0, 1.0])
    ro = np.tile([0.0, 0.0, -3.0], (64, 1))
    a = fwd(scene, ro, rd, use_bvh=False)
    b = fwd(scene, ro, rd, use_bvh=True)
    np.testing.assert_array_equal(a["iters"], b["iters"])
    np.testing.assert_array_equal(a["tri_collection"], b["tri_collection"])
    np.testing.assert_allclose(a["color"], b["color"], rtol=0, atol=0)


def test_primitive_permutation_invariance():
    rng = np.random.default_rng(8)
    n = 50
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    scene = [rng.uniform(-1, 1, (n, 3)), rng.uniform(0.05, 0.3, (n, 3)), q, rng.uniform(0.1, 5, n),
             rng.normal(0, 0.5, (n, 1, 3))]
    perm = rng.permutation(n)
    ro = np.tile([0.0, 0.0, -3.0], (32, 1))
    rd = rng.normal(size=(32, 3)) * 0.2 + np.array([0, 0, 1.0])
    a = fwd(scene, ro, rd)
    b = fwd([x[perm] for x in scene], ro, rd)
    np.testing.assert_allclose(a["color"], b["color"], rtol=1e-12, atol=1e-14)


def test_density_length_scaling_invariance():
    # scaling all lengths by k and densities by 1/k leaves colour and opacity unchanged
    rng = np.random.default_rng(9)
    n = 30
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    means = np.c_[rng.uniform(-0.3, 0.3, (n, 2)), rng.uniform(0.5, 3, n)]
    scales = rng.uniform(0.05, 0.3, (n, 3))
    dens = rng.uniform(0.1, 3, n)
    feats = rng.normal(0, 0.5, (n, 1, 3))
    ro = np.zeros((8, 3))
    rd = np.c_[rng.uniform(-0.1, 0.1, (8, 2)), np.ones(8)]
    a = fwd((means, scales, q, dens, feats), ro, rd)
    k = 2.5
    b = fwd((means * k, scales * k, q, dens / k, feats), ro, rd, tmax=100 * k)
    np.testing.assert_allclose(a["color"][:, :4], b["color"][:, :4], rtol=1e-10, atol=1e-13)


def _fd_scene(seed, n=6, r=5):
    rng = np.random.default_rng(seed)
    means = rng.uniform(-0.5, 0.5, (n, 3))
    means[:, 2] += 1.5
    scales = rng.uniform(0.2, 0.5, (n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    dens = rng.uniform(0.5, 3, n)
    feats = rng.normal(0, 0.5, (n, 1, 3))
    rayo = np.c_[rng.normal(0, 0.02, (r, 2)), np.zeros(r)]
    rayd = np.c_[rng.uniform(-0.2, 0.2, (r, 2)), np.ones(r)]
    g = rng.normal(size=(r, 5))
    return [means, scales, q, dens, feats, rayo, rayd], g


@pytest.mark.parametrize("seed", [1, 2])
def test_backward_matches_central_differences(seed):
    # FD of the safe-variant forward (the function the reference's backward differentiates) vs the adjoint
    params, g = _fd_scene(seed)
    names = ["dmeans", "dscales", "dquats", "ddensities", "dfeatures", "drayo", "drayd"]

    def loss(p):
        f = oracle.forward(*p, 0.0, 100.0, safe=True)
        return float(np.sum(f["color"] * g)), f

    _, f0 = loss(params)
    b = oracle.backward(*params, 0.0, 500, g, f0, clamp=False)
    h = 1e-6
    for i, name in enumerate(names):
        fd = np.zeros_like(params[i])
        for idx in np.ndindex(params[i].shape):
            p = [x.copy() for x in params]
            p[i][idx] += h
            lp, _ = loss(p)
            p[i][idx] -= 2 * h
            lm, _ = loss(p)
            fd[idx] = (lp - lm) / (2 * h)
        scale = max(np.abs(fd).max(), 1e-12)
        assert np.abs(fd - b[name]).max() / scale < 1e-6, name


def test_backward_start_inside_and_tmin_fd():
    # camera inside primitives with tmin > 0: gradients of densities/features flow through the initial state
    rng = np.random.default_rng(11)
    n = 4
    means = np.c_[rng.uniform(-0.05, 0.05, (n, 2)), rng.uniform(0.0, 1.0, n)]
    scales = rng.uniform(0.3, 0.6, (n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    dens = rng.uniform(0.5, 2, n)
    feats = rng.normal(0, 0.5, (n, 1, 3))
    rayo = np.zeros((3, 3))
    rayd = np.c_[rng.uniform(-0.1, 0.1, (3, 2)), np.ones(3)]
    rayd /= np.linalg.norm(rayd, axis=1, keepdims=True)
    g = rng.normal(size=(3, 5))
    tmin = 0.1
    params = [means, scales, q, dens, feats]

    def loss(p):
        f = oracle.forward(*p, rayo, rayd, tmin, 100.0, safe=True)
        return float(np.sum(f["color"] * g)), f

    _, f0 = loss(params)
    assert len(f0["initial_inds"]) > 0
    b = oracle.backward(*params, rayo, rayd, tmin, 500, g, f0, clamp=False)
    h = 1e-6
    for i, name in [(3, "ddensities"), (4, "dfeatures")]:
        fd = np.zeros_like(params[i])
        for idx in np.ndindex(params[i].shape):
            p = [x.copy() for x in params]
            p[i][idx] += h
            lp, _ = loss(p)
            p[i][idx] -= 2 * h
            lm, _ = loss(p)
            fd[idx] = (lp - lm) / (2 * h)
        scale = max(np.abs(fd).max(), 1e-12)
        assert np.abs(fd - b[name]).max() / scale < 1e-6, name


def test_backward_skip_rule():
    # rays with iters >= max_iters - 1 get no gradient (backwards_kernel.slang:304)
    params, g = _fd_scene(3)
    f = oracle.forward(*params, 0.0, 100.0, max_iters=2)
    assert np.all(f["iters"] <= 2)
    b = oracle.backward(*params, 0.0, 2, g, f)
    hit = f["iters"] >= 1
    assert hit.any()
    assert np.all(b["drayo"][hit] == 0) and np.all(b["dmeans"] == 0)


def test_sh_known_answers():
    # sh.slang:18-27 / third_party/sh_util.py:26-54 constants: degree-1 basis at +z is C1 * z on coefficient 2,
    # degree-2 coefficient 6 at +z is C2[2] * 2, DC is C0.
    f = np.zeros((1, 16, 3))
    f[0, 0, 0] = 1.0
    f[0, 2, 1] = 1.0
    f[0, 6, 2] = 1.0
    rgb = oracle.eval_sh(np.array([[0.0, 0.0, 5.0]]), f, 3, np.zeros(3))
    np.testing.assert_allclose(rgb[0], [0.5 + 0.28209479177387814, 0.5 + 0.4886025119029199,
                                        0.5 + 2 * 0.31539156525252005], rtol=1e-12)


def test_sh_basis_orthonormal():
    # the real SH basis (PlenOctree signs) is orthonormal on the sphere: Gram matrix by Lebedev-free quadrature
    nth, nph = 200, 400
    th = (np.arange(nth) + 0.5) * np.pi / nth
    ph = (np.arange(nph) + 0.5) * 2 * np.pi / nph
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    dirs = np.stack([np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)], -1).reshape(-1, 3)
    wq = (np.sin(TH) * (np.pi / nth) * (2 * np.pi / nph)).reshape(-1)
    B = np.zeros((dirs.shape[0], 16))
    for k in range(16):
        f = np.zeros((1, 16, 3))
        f[0, k, 0] = 1.0
        for i0 in range(0, dirs.shape[0], 20000):
            dd = dirs[i0:i0 + 20000]
            # evaluate per direction by placing means along the directions from the origin
            B[i0:i0 + 20000, k] = oracle.eval_sh(dd, np.repeat(f, len(dd), 0), 3, np.zeros(3))[:, 0] - 0.5
    G = (B * wq[:, None]).T @ B
    np.testing.assert_allclose(G, np.eye(16), atol=2e-4)


def test_buffer_rounds_matter_at_a_cutting_tmax():
    # The reference gathers 16 events per traversal (shaders.slang:18, 195-226); an exit beyond tmax is gathered
    # only together with its entry, in the round whose trace reports that entry (:96-106 with OptiX's [tmin, tmax]
    # report interval). So with a tmax that cuts through primitives, the buffer size changes which exits beyond
    # tmax a ray integrates: the oracle, a literal restatement, shows it, and the HIP tracer must follow the
    # 16-event rounds whatever its own buffer size (tests/test_gpu_sweep.py).
    import test_gpu_sweep as S
    from gpu_util import KEYS
    c = S._draw(17)
    sc = [c["sc"][k].double().numpy() for k in KEYS]
    o, d = c["o"].double().numpy(), c["d"].double().numpy()
    a = oracle.forward(*sc, o, d, c["tmin"], c["tmax"], max_iters=c["max_iters"], use_bvh=True, kbuf=16)
    b = oracle.forward(*sc, o, d, c["tmin"], c["tmax"], max_iters=c["max_iters"], use_bvh=True, kbuf=19)
    assert c["tmax"] < 10 and (a["iters"] != b["iters"]).sum() > 0
    # without the cut the buffer size does not matter
    a = oracle.forward(*sc, o, d, c["tmin"], 100.0, max_iters=c["max_iters"], use_bvh=True, kbuf=16)
    b = oracle.forward(*sc, o, d, c["tmin"], 100.0, max_iters=c["max_iters"], use_bvh=True, kbuf=19)
    assert np.array_equal(a["iters"], b["iters"]) and np.abs(a["color"] - b["color"]).max() < 1e-12


def _real_sh_scipy(dirs):
    """The 16 real SH basis functions of degrees 0-3 at unit directions, from scipy's complex Y_l^m (which carries
    the Condon-Shortley phase): index l*l + l + m holds Y_l^0 for m = 0, sqrt(2) Re Y_l^m for m > 0 and
    sqrt(2) Im Y_l^|m| for m < 0. Keeping the phase is what gives the PlenOctree signs of sh.slang:48-88 (for
    example -C1*y, C1*z, -C1*x at degree 1)."""
    from scipy.special import sph_harm_y
    th = np.arccos(np.clip(dirs[:, 2], -1.0, 1.0))
    ph = np.arctan2(dirs[:, 1], dirs[:, 0])
    S = np.zeros((dirs.shape[0], 16))
    for l in range(4):
        for m in range(-l, l + 1):
            Y = sph_harm_y(l, abs(m), th, ph)
            S[:, l * l + l + m] = Y.real if m == 0 else np.sqrt(2) * (Y.imag if m < 0 else Y.real)
    return S


def _oracle_sh_basis(dirs):
    # basis function k = colour of a one-hot feature on coefficient k, minus the 0.5 offset (sh.slang:48-51)
    B = np.zeros((dirs.shape[0], 16))
    for k in range(16):
        f = np.zeros((dirs.shape[0], 16, 3))
        f[:, k, 0] = 1.0
        B[:, k] = oracle.eval_sh(dirs, f, 3, np.zeros(3))[:, 0] - 0.5
    return B


def test_sh_basis_matches_scipy_real_harmonics():
    # all 16 basis functions of sh.slang:48-88 (constants third_party/sh_util.py:26-54) against scipy at generic
    # directions; a flipped sign or two coefficients swapped within a degree must fail this comparison
    rng = np.random.default_rng(11)
    v = rng.normal(size=(32, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    B, S = _oracle_sh_basis(v), _real_sh_scipy(v)
    np.testing.assert_allclose(B, S, rtol=0, atol=1e-13)
    # the comparison separates every such mistake: each single sign flip and each within-degree swap moves some
    # basis value by far more than the tolerance
    for k in range(1, 16):
        assert np.abs(-S[:, k] - S[:, k]).max() > 1e-2
    for l in range(1, 4):
        for i in range(l * l, (l + 1) ** 2):
            for j in range(i + 1, (l + 1) ** 2):
                assert np.abs(S[:, i] - S[:, j]).max() > 1e-2


def _means2d_scene(seed=3, n=8, r=12):
    rng = np.random.default_rng(seed)
    means = rng.uniform(-0.4, 0.4, (n, 3))
    scales = rng.uniform(0.15, 0.4, (n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    dens = rng.uniform(0.5, 3, n)
    feats = rng.normal(0, 0.5, (n, 1, 3))
    eye = np.array([0.3, -0.2, -3.0])
    rayo = np.repeat(eye[None], r, 0) + rng.normal(0, 0.01, (r, 3))
    rayd = np.c_[rng.uniform(-0.15, 0.15, (r, 2)), np.ones(r)]
    g = rng.normal(size=(r, 5))
    return [means, scales, q, dens, feats, rayo, rayd], g


def _wct(rng):
    # world-to-camera in the row-vector convention of backwards_kernel.slang:244-250 (p_view = [x, y, z, 1] @ W)
    a = rng.normal(size=(3, 3))
    Rm, _ = np.linalg.qr(a)
    W = np.eye(4)
    W[:3, :3] = Rm
    W[3, :3] = rng.normal(0, 0.3, 3) + np.array([0.1, -0.2, 3.5])  # translation keeps the scene in front
    return W


def _inv_project_vjp(mean, W, upstream, transpose=False):
    """bwd_diff(inv_project) w.r.t. xy (backwards_kernel.slang:213-220, 244-257) by float64 torch autograd: xy and
    dist come from project(mean, W); inv_project(xy, dist, inv(W)) = ([xy * dist, dist, 1] @ inv(W))[:3]."""
    import torch
    Wt = torch.tensor(W, dtype=torch.float64)
    m = torch.tensor(mean, dtype=torch.float64)
    pv = torch.cat([m, torch.ones(1, dtype=torch.float64)]) @ Wt
    xy = (pv[:2] / pv[2]).detach().requires_grad_(True)
    dist = pv[2].detach()
    inv = torch.linalg.inv(Wt)
    if transpose:
        inv = inv.T
    p_hom = torch.cat([xy * dist, dist.reshape(1), torch.ones(1, dtype=torch.float64)])
    out = (p_hom @ inv)[:3]
    out.backward(torch.tensor(upstream, dtype=torch.float64))
    return xy.grad.numpy()


def test_means2d_statistic_one_matrix_matches_autograd():
    params, g = _means2d_scene()
    W = _wct(np.random.default_rng(4))
    f = oracle.forward(*params, 0.0, 100.0)
    b = oracle.backward(*params, 0.0, 500, g, f, wcts=W[None], clamp=False)
    assert np.abs(b["dmeans"]).max() > 1e-3
    # one matrix: every event of primitive p has the same Jacobian, so the statistic is its transpose applied to
    # the summed dL/dmean
    exp = np.stack([_inv_project_vjp(params[0][p], W, b["dmeans"][p]) for p in range(params[0].shape[0])])
    np.testing.assert_allclose(b["dmeans2D"], exp, rtol=1e-10, atol=1e-12)
    wrong = np.stack([_inv_project_vjp(params[0][p], W, b["dmeans"][p], transpose=True)
                      for p in range(params[0].shape[0])])
    assert np.abs(wrong - exp).max() > 1e-2 * np.abs(exp).max()


def test_means2d_statistic_per_ray_matrices_matches_autograd():
    params, g = _means2d_scene(seed=5)
    rng = np.random.default_rng(6)
    nr = params[5].shape[0]
    Ws = np.stack([_wct(rng) for _ in range(nr)])
    f = oracle.forward(*params, 0.0, 100.0)
    b = oracle.backward(*params, 0.0, 500, g, f, wcts=Ws, clamp=False)
    n = params[0].shape[0]
    exp = np.zeros((n, 2))
    wrong = np.zeros((n, 2))
    for r in range(nr):
        # the dL/dmean that ray r alone contributes, and its Jacobian through ray r's matrix
        gr = np.zeros_like(g)
        gr[r] = g[r]
        br = oracle.backward(*params, 0.0, 500, gr, f, clamp=False)
        for p in range(n):
            if np.any(br["dmeans"][p] != 0):
                exp[p] += _inv_project_vjp(params[0][p], Ws[r], br["dmeans"][p])
                wrong[p] += _inv_project_vjp(params[0][p], Ws[r], br["dmeans"][p], transpose=True)
    assert np.abs(exp).max() > 1e-3
    np.testing.assert_allclose(b["dmeans2D"], exp, rtol=1e-9, atol=1e-12)
    assert np.abs(wrong - exp).max() > 1e-2 * np.abs(exp).max()
