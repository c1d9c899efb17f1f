"""Mirror of the reference's eval_sh.py (EvalSH autograd Function, eval_sh) on the HIP library.

eval_sh.py:28-82: per-primitive colour SH(features, normalize(mean - rayo)) + 0.5, unclamped; backward gives
dL/dfeatures only (sh_kernel.slang:45-98).
"""
from typing import Any

import torch
from torch.autograd import Function

from . import _capi as C


class EvalSH(Function):
    @staticmethod
    def forward(ctx: Any, means: torch.Tensor, features: torch.Tensor, rayo: torch.Tensor, sh_degree: int):
        rayo = rayo.reshape(3).contiguous().float()
        means = means.contiguous()
        features = features.contiguous()
        color = torch.empty_like(means)
        ctx.sh_degree = sh_degree
        with torch.cuda.device(means.device):
            C.check(C.lib().ever_eval_sh(C.ptr(means), C.ptr(features), C.ptr(rayo), means.shape[0],
                                         features.shape[1], int(sh_degree), C.ptr(color), C.stream_ptr()),
                    "ever_eval_sh")
        ctx.save_for_backward(means, features, rayo, color)
        return color

    @staticmethod
    def backward(ctx, dL_dcolor: torch.Tensor):
        means, features, rayo, color = ctx.saved_tensors
        dL_dfeat = torch.empty_like(features)
        dL_dcolor = dL_dcolor.contiguous()
        with torch.cuda.device(means.device):
            C.check(C.lib().ever_eval_sh_backward(C.ptr(means), C.ptr(features), C.ptr(rayo), C.ptr(dL_dcolor),
                                                  means.shape[0], features.shape[1], int(ctx.sh_degree),
                                                  C.ptr(dL_dfeat), C.stream_ptr()),
                    "ever_eval_sh_backward")
        return None, dL_dfeat, None, None


def eval_sh(means, features, rayo, sh_degree):
    return EvalSH.apply(means, features, rayo, sh_degree)
