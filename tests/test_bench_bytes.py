"""The bench's algorithmic byte count (bench.algorithmic_bytes) on the CPU: a hand-counted event list, and the
same events stored by processing slot (columns permuted by ray_order) giving the same counts."""
import os
import sys

import torch

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def _events():
    # 3 rays, max_iters 4; ids 2*prim + entry: ray 0 enters and leaves prim 5, ray 1 enters 1 and 2 and leaves 1,
    # ray 2 has no event
    max_iters, nr = 4, 3
    tri = torch.zeros((max_iters, nr), dtype=torch.int32)
    tri[0, 0], tri[1, 0] = 2 * 5 + 1, 2 * 5
    tri[0, 1], tri[1, 1], tri[2, 1] = 2 * 1 + 1, 2 * 2 + 1, 2 * 1
    iters = torch.tensor([2, 3, 0], dtype=torch.int32)
    return tri, iters, max_iters, nr


def test_counts_by_hand():
    tri, iters, max_iters, nr = _events()
    out = dict(iters=iters, tri_collection=tri.reshape(-1))
    ab = bench.algorithmic_bytes(out, nr, 10, 16, max_iters)
    assert ab["events"] == 5
    assert ab["distinct"] == 1 + 2  # max(entries, exits) per ray: 1, 2, 0
    assert ab["fwd"] == nr * 44 + 3 * bench.P_BYTES + nr * 68 + 4 * 5


def test_by_slot_layout_gives_the_same_counts():
    tri, iters, max_iters, nr = _events()
    order = torch.tensor([2, 0, 1], dtype=torch.int32)  # slot s traces ray order[s]
    by_slot = tri[:, order.long()].contiguous()
    a = bench.algorithmic_bytes(dict(iters=iters, tri_collection=tri.reshape(-1)), nr, 10, 16, max_iters)
    b = bench.algorithmic_bytes(dict(iters=iters, tri_collection=by_slot.reshape(-1), ray_order=order,
                                     event_layout="slot"), nr, 10, 16, max_iters)
    assert a == b
