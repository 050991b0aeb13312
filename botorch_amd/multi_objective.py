"""The fused log-areas op of qLogEHVI / qLogNEHVI on libbotorch_amd.

Mirrors ``_FusedLogAreas`` (acquisition/multi_objective/logei.py:107-170), the
autograd Function around the reference's only native op (csrc/logei_fused.cpp):
steps 1-4 of ``_compute_log_qehvi`` (logei.py:437-479) - log smoothed
improvement over the cell lower bounds, smooth minimum over the subset members,
smooth minimum with the log cell side lengths, sum over objectives - for
objective subsets ``batch x n_subsets x i x m`` and cell bounds ``num_cells x m``
(or one set per batch element), giving ``batch x num_cells x n_subsets``.  Here
both directions run as HIP kernels (``bta_log_areas_forward`` / ``_backward``)
on the tensors' device, in their dtype (float32 or float64).
"""
from __future__ import annotations

import math

import torch
from torch import Tensor

from . import _lib

TAU_RELU = 1e-6  # acquisition/logei.py:67
TAU_MAX = 1e-2  # acquisition/logei.py:68
FUSED_MAX_I = 32  # subset size bound of the op (logei.py:62)
FUSED_MAX_M = 8  # objective bound of the op (csrc/logei_fused.cpp:33)


def _flat(obj_subsets: Tensor, cell_lower: Tensor, cell_upper: Tensor):
    if obj_subsets.dtype not in (torch.float32, torch.float64):
        raise TypeError("obj_subsets must be float32 or float64")
    if cell_lower.dtype != obj_subsets.dtype or cell_upper.dtype != obj_subsets.dtype:
        raise TypeError("cell bounds must have the dtype of obj_subsets")
    if obj_subsets.dim() < 3 or cell_lower.dim() < 2 or cell_lower.shape != cell_upper.shape:
        raise ValueError("obj_subsets must be ... x n_subsets x i x m and the cell bounds "
                         "... x num_cells x m")
    n_sub, i, m = obj_subsets.shape[-3:]
    batch_shape = obj_subsets.shape[:-3]
    B = max(1, math.prod(batch_shape))
    nc = cell_lower.shape[-2]
    if cell_lower.shape[-1] != m or i > FUSED_MAX_I or m > FUSED_MAX_M:
        raise ValueError(f"need matching m <= {FUSED_MAX_M} and i <= {FUSED_MAX_I}")
    obj = obj_subsets.reshape(B, n_sub, i, m).contiguous()
    cl = cell_lower.reshape(-1, nc, m).contiguous()
    cu = cell_upper.reshape(-1, nc, m).contiguous()
    if cl.shape[0] not in (1, B):
        raise ValueError("batched cell bounds must have one set per batch element")
    return obj, cl, cu, batch_shape, B, n_sub, i, m, nc


class _FusedLogAreas(torch.autograd.Function):
    @staticmethod
    def forward(ctx, obj_subsets: Tensor, cell_lower: Tensor, cell_upper: Tensor,
                tau_relu: float, tau_max: float) -> Tensor:
        obj, cl, cu, batch_shape, B, n_sub, i, m, nc = _flat(obj_subsets, cell_lower,
                                                             cell_upper)
        out = torch.empty(B, nc, n_sub, dtype=obj.dtype, device=obj.device)
        f32 = int(obj.dtype == torch.float32)
        _lib.check(_lib.lib().bta_log_areas_forward(
            f32, _lib.ptr(obj), B, n_sub, i, m, _lib.ptr(cl), _lib.ptr(cu), cl.shape[0], nc,
            float(tau_relu), float(tau_max), _lib.ptr(out), _lib.stream()), "log_areas_forward")
        ctx.save_for_backward(obj, cl, cu)
        ctx.tau = (float(tau_relu), float(tau_max))
        ctx.orig_shape = obj_subsets.shape
        return out.reshape(*batch_shape, nc, n_sub)

    @staticmethod
    def backward(ctx, grad_output: Tensor):
        obj, cl, cu = ctx.saved_tensors
        B, n_sub, i, m = obj.shape
        nc = cl.shape[1]
        go = grad_output.reshape(B, nc, n_sub).to(obj.dtype).contiguous()
        g_obj = torch.empty_like(obj)
        f32 = int(obj.dtype == torch.float32)
        _lib.check(_lib.lib().bta_log_areas_backward(
            f32, _lib.ptr(go), _lib.ptr(obj), B, n_sub, i, m, _lib.ptr(cl), _lib.ptr(cu),
            cl.shape[0], nc, ctx.tau[0], ctx.tau[1], _lib.ptr(g_obj), _lib.stream()),
            "log_areas_backward")
        return g_obj.reshape(ctx.orig_shape), None, None, None, None


def fused_log_areas(obj_subsets: Tensor, cell_lower: Tensor, cell_upper: Tensor,
                    tau_relu: float = TAU_RELU, tau_max: float = TAU_MAX) -> Tensor:
    """Log areas ``batch x num_cells x n_subsets`` of the objective subsets in the
    hyper-rectangle cells, differentiable in ``obj_subsets``."""
    return _FusedLogAreas.apply(obj_subsets, cell_lower, cell_upper, tau_relu, tau_max)
