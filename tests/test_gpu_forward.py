"""Forward parity of the HIP tracer against the FP64 oracle (colour and opacity within 1e-4)."""
import numpy as np
import pytest
import torch

from ever_amd import scenes
from gpu_util import gpu_forward, oracle_forward

pytestmark = pytest.mark.gpu


def test_config1_forward_matches_oracle():
    sc, o, d = scenes.config1()
    print("case: config1 forward", flush=True)
    _, out = gpu_forward(sc, o, d)
    torch.cuda.synchronize()
    ref = oracle_forward(sc, o, d)
    col = out["color"].cpu().numpy()
    np.testing.assert_allclose(col[:, :4], ref["color"][:, :4], atol=1e-4, rtol=1e-4)
    np.testing.assert_allclose(col[:, 4], ref["color"][:, 4], atol=1e-3, rtol=1e-3)
    it = out["iters"].cpu().numpy()
    assert (np.abs(it - ref["iters"]) <= 1).mean() > 0.999
