"""Scene I/O on the GPU against the PLY oracle: a 3D Gaussian Splatting file through the viewer's activations
(ply_file_loader.cc:164-193), a save_ply file read back bit for bit, Morton orders equal to the oracle's (exact
integer work, ties and degenerate bounds included), Morton-ordered loading, and the image of a reordered scene."""
import numpy as np
import pytest
import torch

from ever_amd import ply as P
from ever_amd import scenes
from ever_amd import tracer as T
from oracle import ply as O

pytestmark = pytest.mark.gpu
D = torch.device("cuda:0")


def _write_3dgs(path, n, seed, num_rest=15):
    """A file in the 3DGS writer's property order with raw (pre-activation) values."""
    rng = np.random.default_rng(seed)
    names = ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
    names += [f"f_rest_{i}" for i in range(3 * num_rest)] + ["opacity", "scale_0", "scale_1", "scale_2"]
    names += ["rot_0", "rot_1", "rot_2", "rot_3"]
    rec = rng.normal(size=(n, len(names))).astype(np.float32)
    rec[:, 3:6] = 0
    rec[:, -8] *= 4  # opacity logits
    rec[:, -7:-4] = rng.uniform(-8, 8, (n, 3))  # raw scales: both softplus branches and the cap
    rec[:, -7] += 30 * (rng.random(n) < 0.05)
    head = "ply\nformat binary_little_endian 1.0\nelement vertex {}\n{}end_header\n".format(
        n, "".join(f"property float {nm}\n" for nm in names))
    with open(path, "wb") as f:
        f.write(head.encode())
        f.write(rec.astype("<f4").tobytes())
    return names, rec


@pytest.mark.parametrize("num_rest", [15, 3, 0])
def test_3dgs_file_gets_the_viewer_activations(tmp_path, num_rest):
    print(f"case: 3dgs file, num_rest {num_rest}", flush=True)
    path = str(tmp_path / "g.ply")
    names, rec = _write_3dgs(path, 5000, 1 + num_rest, num_rest)
    out = P.load_scene(path, D)
    col = {nm: i for i, nm in enumerate(names)}
    sc, dens, q = O.activate_3dgs(rec[:, [col[f"scale_{a}"] for a in range(3)]], rec[:, col["opacity"]],
                                  rec[:, [col[f"rot_{a}"] for a in range(4)]])
    rel = lambda a, b: float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))
    assert np.array_equal(out["means"].cpu().numpy(), rec[:, 0:3])
    # the loader's float sums (1 + exp(s) in its softplus, 1 - alpha) round to one unit of 1.0 (1.2e-7) apart
    # when expf/logf differ by an ulp, so the bounds carry that absolute part; elsewhere a few float ulps
    gs = out["scales"].cpu().numpy().astype(np.float64)
    assert (np.abs(gs - sc) <= 1.5e-7 + 1e-6 * sc).all()
    ms = np.min(2 * sc, axis=1)
    bound = dens * (2e-6 + 3e-7 / ms) + 1.5e-7 / ms
    assert (np.abs(out["densities"].cpu().numpy() - dens) <= bound).all()
    assert rel(out["quats"].cpu().numpy(), q) < 1e-6
    f = out["features"].cpu().numpy()
    assert f.shape == (5000, 1 + num_rest, 3)
    assert np.array_equal(f[:, 0], rec[:, 6:9])
    r = rec[:, 9:9 + 3 * num_rest].reshape(5000, 3, num_rest)
    assert np.array_equal(f[:, 1:], r.transpose(0, 2, 1))


def test_save_ply_round_trip_is_exact(tmp_path):
    sc = scenes.clustered_scene(3000, 3, seed=4)
    path = str(tmp_path / "s.ply")
    P.save_ply(sc["means"], sc["scales"], sc["quats"], sc["densities"], sc["features"], path)
    out = P.load_scene(path, D)
    for k in ("means", "scales", "quats", "densities", "features"):
        assert torch.equal(out[k].cpu(), sc[k].reshape(out[k].shape)), k


@pytest.mark.parametrize("case", ["uniform", "ties", "negative", "single", "empty", "large"])
def test_morton_order_matches_oracle(case):
    print(f"case: morton {case}", flush=True)
    rng = np.random.default_rng(5)
    if case == "uniform":
        m = rng.uniform(-2, 2, (20000, 3))
    elif case == "ties":
        m = rng.integers(0, 4, (5000, 3)).astype(np.float64)  # many equal codes
    elif case == "negative":
        m = -rng.uniform(1, 2, (3000, 3))
    elif case == "single":
        m = np.array([[0.5, -0.25, 3.0]])
    elif case == "empty":
        m = np.zeros((0, 3))
    else:
        m = rng.normal(0, 1, (1_500_000, 3))
    m = m.astype(np.float32)
    got = P.morton_order(torch.from_numpy(m).to(D).contiguous()).cpu().numpy()
    want = O.morton_order(m)
    assert got.dtype == np.int32 and np.array_equal(got, want)


def test_morton_loading_and_trace_invariance(tmp_path):
    sc = scenes.uniform_scene(4000, 3, seed=6)
    path = str(tmp_path / "u.ply")
    P.save_ply(sc["means"], sc["scales"], sc["quats"], sc["densities"], sc["features"], path)
    a = P.load_scene(path, D)
    b = P.load_scene(path, D, morton=True)
    order = b["order"].long()
    assert np.array_equal(b["order"].cpu().numpy(), O.morton_order(sc["means"].numpy()))
    for k in ("means", "scales", "quats", "densities", "features"):
        assert torch.equal(b[k], a[k][order]), k
    o, d = scenes.pinhole_rays(96, 64, (0.0, 0.0, -3.0))
    o, d = o.to(D).contiguous(), d.to(D).contiguous()
    ca = T.trace_forward(T.Scene(a["means"], a["scales"], a["quats"], a["densities"], a["features"]).build(),
                         o, d, 0.0, 100.0, 500)["color"]
    cb = T.trace_forward(T.Scene(b["means"], b["scales"], b["quats"], b["densities"], b["features"]).build(),
                         o, d, 0.0, 100.0, 500)["color"]
    assert float((ca - cb).abs().max()) < 1e-5
