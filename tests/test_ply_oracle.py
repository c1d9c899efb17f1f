"""Pins the PLY-preparation oracle (oracle/ply.py) on the CPU by closed forms of the viewer's activations
(ply_file_loader.cc:150-193) and by properties of the Morton codes (:72-148) that do not re-evaluate them, and
checks the host side of ever_amd.ply (header parsing and its errors, the save_ply layout of
utils/math_util.py:152-202) without a GPU."""
import math

import numpy as np
import pytest

from ever_amd import ply as P
from oracle import ply as O


def test_activations_closed_forms():
    s = np.array([[0.0, 10.0, 30.0]])
    q = np.array([[2.0, 0.0, 0.0, 0.0]])
    scales, dens, quats = O.activate_3dgs(s, np.array([60.0]), q)
    assert np.allclose(scales, [[math.log(2) + 3e-4, 10 + 3e-4, 25.0]], rtol=1e-7, atol=0)
    # opacity -> +inf: alpha = 0.99, density = -log(0.01) / min(2 s)
    assert abs(dens[0] / ((-math.log(0.01)) / (2 * (math.log(2) + 3e-4))) - 1) < 3e-7
    # the float sum 1 + exp(s) of the loader's softplus: for s = -12 it keeps about one significant digit
    e = math.exp(-12.0)
    sc12, _, _ = O.activate_3dgs(np.array([[-12.0, 0.0, 0.0]]), np.array([0.0]), q)
    f1 = float(np.float32(1.0 + e))
    assert f1 != 1.0 + e and abs(sc12[0, 0] - float(np.float32(float(np.float32(math.log(f1))) + 3e-4))) < 1e-12
    assert np.allclose(quats, [[0.5, 0, 0, 0]])  # divided by the squared length
    _, d0, _ = O.activate_3dgs(s, np.array([-60.0]), q)
    assert 0 <= d0[0] < 1e-25


def _deinterleave(code, a):
    return sum(((int(code) >> (3 * b + a)) & 1) << b for b in range(21))


def test_morton_codes_decode_to_quantised_coordinates():
    rng = np.random.default_rng(1)
    m = rng.uniform(-3, 2, (500, 3)).astype(np.float32)
    codes = O.morton_codes(m)
    lo, hi = m.min(0).astype(np.float64), m.max(0).astype(np.float64)
    for i in range(0, 500, 7):
        for a in range(3):
            want = (m[i, a] - lo[a]) / (hi[a] - lo[a]) * 2097151
            assert abs(_deinterleave(codes[i], a) - want) <= 2.0  # float32 rounding, then truncation
    # corners and axis extremes
    c = O.morton_codes(np.array([[0, 0, 0], [1, 1, 1], [1, 0, 0], [0, 0, 1]], np.float32))
    assert c[0] == 0 and c[1] == 2 ** 63 - 1
    assert c[2] == sum(8 ** b for b in range(21)) and c[3] == 4 * sum(8 ** b for b in range(21))


def test_morton_order_is_the_sorted_order():
    rng = np.random.default_rng(2)
    m = rng.uniform(-1, 1, (300, 3)).astype(np.float32)
    m[50:60] = m[10]  # ties keep index order
    order = O.morton_order(m)
    codes = O.morton_codes(m)
    brute = sorted(range(300), key=lambda i: (int(codes[i]), i))
    assert order.tolist() == brute


def test_morton_degenerate_bounds():
    assert O.morton_codes(np.array([[0.3, 0.3, 0.3]], np.float32))[0] == 0  # zero extent
    # all coordinates negative: the maximum starts at the smallest positive float, as the loader's does
    m = np.array([[-2.0, -1.0, -3.0], [-1.0, -2.0, -1.0]], np.float32)
    c = O.morton_codes(m)
    assert _deinterleave(c[1], 0) == int(np.float32(2097151) * np.float32((m[1, 0] - m[0, 0]) / (O.FLT_MIN - m[0, 0])))


def test_save_ply_layout_and_read_back(tmp_path):
    rng = np.random.default_rng(3)
    n = 17
    f = rng.normal(size=(n, 16, 3)).astype(np.float32)
    means, sc = rng.normal(size=(n, 3)).astype(np.float32), rng.random((n, 3)).astype(np.float32)
    q, d = rng.normal(size=(n, 4)).astype(np.float32), rng.random(n).astype(np.float32)
    path = str(tmp_path / "a" / "s.ply")
    P.save_ply(means, sc, q, d, f, path)
    names, rec = P.read_ply(path)
    assert names[:9] == ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
    assert names[9:54] == [f"f_rest_{i}" for i in range(45)]
    assert names[54:] == ["density", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]
    assert np.array_equal(rec[:, :3], means) and not rec[:, 3:6].any()
    assert np.array_equal(rec[:, 6:9], f[:, 0])
    # f_rest channel-major: f_rest[c * 15 + k] = features[:, 1 + k, c]
    assert np.array_equal(rec[:, 9:54].reshape(n, 3, 15), f[:, 1:].transpose(0, 2, 1))
    assert np.array_equal(rec[:, 54], d) and np.array_equal(rec[:, 55:58], sc) and np.array_equal(rec[:, 58:], q)


def test_header_errors(tmp_path):
    p = tmp_path / "bad.ply"
    for text, msg in ((b"plx\n", '"ply" header'), (b"ply\nformat ascii 1.0\n", "binary_little_endian"),
                      (b"ply\nformat binary_little_endian 1.0\nelement face 3\n", "element vertex"),
                      (b"ply\nformat binary_little_endian 1.0\nelement vertex 2\nproperty float x\nend_header\n\0\0",
                       "truncated"),
                      (b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty uchar x\nend_header\n\0",
                       "only float")):
        p.write_bytes(text)
        with pytest.raises(ValueError, match=msg):
            P.read_ply(str(p))
    names = ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2",
             "rot_3"]
    with pytest.raises(ValueError, match="density"):
        P._layout(names, None)
    L = P._layout(names + ["opacity"], None)
    assert L.activation == 1 and L.opacity == len(names) and L.num_rest == 0
