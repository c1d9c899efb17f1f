"""Degenerate and deep hierarchies: the traversal stacks hold every entry a tree can need (at most 180: a
root-to-leaf path of the Karras hierarchy has at most 60 internal nodes, 30 split levels of the 30-bit keys and 30
of the leaf index, and a 4-wide record pushes at most three children), so no push is ever dropped. OptiX, which
the LBVH replaces, puts no depth limit on the caller (reference: splinetracers/fast_ellipsoid_splinetracer/
GAS.cpp:56-153). Each scene is traced by both walks and checked against the FP64 oracle, and the stack need of the
built records is measured from the scene workspace."""
import numpy as np
import pytest
import torch

from ever_amd import _capi as C
from ever_amd import scenes
from ever_amd import tracer as T
from gpu_util import KEYS, dev, oracle_forward

pytestmark = pytest.mark.gpu

EMPTY = -2 ** 31


def record_stack_need(scene):
    """(records on the deepest path, largest stack a walk can build) of the built 4-wide records: a record pushes
    all its children but the first it enters, so need(i) = children(i) - 1 + max(0, need of its record children)."""
    lay = C.scene_layout(scene.n)
    ws = scene.workspace.cpu().numpy()
    nleaves = int(ws[lay["header"]:lay["header"] + 64].view(np.uint32)[9])
    if nleaves <= 1:
        return 0, 0
    rec = ws[lay["nodes4"]:lay["nodes4"] + 128 * (nleaves - 1)].view(np.int32).reshape(-1, 32)
    refs = rec[:, 24:28]
    depth, need = {}, {}
    stack = [(0, False)]
    while stack:
        i, done = stack.pop()
        kids = [int(c) for c in refs[i] if c != EMPTY]
        inner = [c for c in kids if c >= 0]
        if not done:
            stack.append((i, True))
            stack.extend((c, False) for c in inner)
            continue
        depth[i] = 1 + max((depth[c] for c in inner), default=0)
        need[i] = len(kids) - 1 + max((need[c] for c in inner), default=0)
    return depth[0], need[0]


def _check_both_walks(sc, o, d, max_iters, sub):
    D = dev()
    t = [sc[k].to(D).contiguous() for k in KEYS]
    scene = T.Scene(*t).build()
    depth, need = record_stack_need(scene)
    print(f"  records on the deepest path {depth}, stack need {need}", flush=True)
    assert need <= 180
    ref = oracle_forward(sc, o[sub], d[sub], max_iters=max_iters, use_bvh=True)
    od, dd = o.to(D).contiguous(), d.to(D).contiguous()
    outs = [T.trace_forward(scene, od, dd, 0.0, 100.0, max_iters, traversal=w) for w in ("ray", "wave")]
    for out in outs:
        col = out["color"].cpu().numpy()[sub]
        assert np.abs(col[:, :4] - ref["color"][:, :4]).max() <= 1e-4
        assert np.array_equal(out["iters"].cpu().numpy()[sub], ref["iters"])
    for k in ("color", "iters", "states"):
        assert torch.equal(outs[0][k], outs[1][k]), k
    return depth, need


def test_coincident_cluster_with_far_outlier():
    # 200,000 primitives whose centres share one Morton cell (the far outlier stretches the centroid bounds), so
    # the hierarchy below that cell splits on the leaf index alone
    print("case: 200k near-coincident primitives + one far outlier", flush=True)
    n = 200_000
    g = torch.Generator().manual_seed(31)
    means = torch.randn((n, 3), generator=g) * 1e-3
    means[-1] = torch.tensor([60.0, 55.0, 50.0])
    sc = dict(means=means, scales=torch.full((n, 3), 0.01) * (1 + torch.rand((n, 3), generator=g)),
              quats=scenes._unit_quats(g, n), densities=torch.full((n,), 0.5),
              features=torch.randn((n, 1, 3), generator=g) * 0.5)
    sc["scales"][-1] = torch.tensor([2.0, 2.0, 2.0])
    o, d = scenes.pinhole_rays(96, 96, (0.0, 0.0, -0.3), fov_deg=20.0)
    sub = np.arange(0, o.shape[0], 37)
    depth, _ = _check_both_walks(sc, o, d, 48, sub)
    assert depth >= 9


def test_morton_comb_over_a_dense_cell():
    # one primitive splitting off at every one of the 27 Morton levels towards a corner cell that holds 2^17
    # primitives of eight size classes: Morton levels, size-class levels and index levels on one path
    print("case: Morton comb + 131072-primitive cell", flush=True)
    g = torch.Generator().manual_seed(32)
    ext, cells = 4.0, 512
    comb = []
    for level in range(27):
        cell = [0, 0, 0]
        cell[level % 3] = 1 << (8 - level // 3)
        comb.append([(c + 0.5) / cells * ext for c in cell])
    comb.append([(cells - 0.5) / cells * ext] * 3)
    comb = torch.tensor(comb)
    m = 1 << 17
    dense = torch.rand((m, 3), generator=g) * (0.9 * ext / cells)
    means = torch.cat([dense, comb])
    n = means.shape[0]
    # dense-cell sizes spread over the size classes (one per halving of the box relative to the scene)
    size = 2.0 ** -torch.randint(3, 11, (m, 1), generator=g).float() * ext
    scales = torch.cat([size.repeat(1, 3) * (0.5 + 0.5 * torch.rand((m, 3), generator=g)),
                        torch.full((comb.shape[0], 3), 0.3 * ext / cells)])
    sc = dict(means=means, scales=scales, quats=scenes._unit_quats(g, n),
              densities=torch.exp(torch.empty(n).uniform_(-2.0, 1.0, generator=g)),
              features=torch.randn((n, 1, 3), generator=g) * 0.5)
    eye = (-0.6, -0.5, -0.4)
    o, d = scenes.pinhole_rays(128, 128, eye, fov_deg=50.0, target=(0.02, 0.02, 0.02))
    sub = np.arange(0, o.shape[0], 53)
    depth, _ = _check_both_walks(sc, o, d, 40, sub)
    assert depth >= 12
