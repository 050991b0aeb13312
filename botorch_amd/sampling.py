"""Sobol-QMC normal base samples (sampling/normal.py:173-213, sampling/qmc.py:59-97,
utils/sampling.py:51-143).

Base samples come from ``torch.quasirandom.SobolEngine`` - the same engine the
reference uses - followed by the inverse-normal transform with the eps clamp of
sampling/qmc.py:80-83, so they are bit-identical to the reference's for the
same seed.  They are generated once per shape (as in the reference) and moved
to the GPU; the per-step sampling (mean + L z) runs in libbotorch_amd.
"""
from __future__ import annotations

import math

import torch
from torch import Tensor
from torch.quasirandom import SobolEngine

from .exceptions import UnsupportedError


def draw_sobol_normal_samples(d: int, n: int, device=None, dtype=torch.float64,
                              seed: int | None = None) -> Tensor:
    """n x d quasi-normal samples (utils/sampling.py:114-143, qmc.py:80-83)."""
    eng = SobolEngine(dimension=d, scramble=True, seed=seed)
    u = eng.draw(n, dtype=dtype)
    v = 0.5 + (1 - torch.finfo(u.dtype).eps) * (u - 0.5)
    return (torch.erfinv(2 * v - 1) * math.sqrt(2)).to(device=device)


def draw_sobol_samples(bounds: Tensor, n: int, q: int, batch_shape=None,
                       seed: int | None = None) -> Tensor:
    """n x q x d scaled Sobol points in the box ``bounds`` (utils/sampling.py:51-99)."""
    batch_shape = batch_shape or torch.Size()
    batch_size = int(torch.Size(batch_shape).numel())
    d = bounds.shape[-1]
    eng = SobolEngine(q * d, scramble=True, seed=seed)
    raw = eng.draw(batch_size * n, dtype=bounds.dtype)
    raw = raw.view(*batch_shape, n, q, d).to(device=bounds.device)
    if batch_shape != torch.Size():
        raw = raw.permute(-3, *range(len(batch_shape)), -2, -1)
    lo, hi = bounds[0], bounds[1]
    return lo + (hi - lo) * raw


class SobolQMCNormalSampler(torch.nn.Module):
    """Sampler with t-batch-collapsed Sobol base samples (sampling/base.py:32-150)."""

    def __init__(self, sample_shape: torch.Size, seed: int | None = None):
        super().__init__()
        if not isinstance(sample_shape, torch.Size):
            raise UnsupportedError("sample_shape must be a torch.Size")
        self.sample_shape = sample_shape
        self.seed = seed if seed is not None else int(torch.randint(0, 1000000, (1,)).item())
        self.register_buffer("base_samples", None)

    def _get_collapsed_shape(self, posterior) -> torch.Size:
        bss = posterior.base_sample_shape
        start, end = posterior.batch_range
        end = len(bss) + end if end < 0 else end
        return self.sample_shape + torch.Size(
            [1 if start <= i < end else s for i, s in enumerate(bss)])

    def _construct_base_samples(self, posterior) -> None:
        target = self._get_collapsed_shape(posterior)
        if self.base_samples is None or self.base_samples.shape != target:
            out_dim = target[len(self.sample_shape):].numel()
            if out_dim > SobolEngine.MAXDIM:
                raise UnsupportedError(
                    f"SobolQMCSampler only supports dimensions q * o <= {SobolEngine.MAXDIM}")
            bs = draw_sobol_normal_samples(out_dim, self.sample_shape.numel(),
                                           device=posterior.device, dtype=posterior.dtype,
                                           seed=self.seed)
            self.base_samples = bs.view(target)

    def forward(self, posterior) -> Tensor:
        self._construct_base_samples(posterior)
        return posterior.rsample_from_base_samples(
            self.sample_shape,
            self.base_samples.expand(self.sample_shape + posterior.base_sample_shape))


def get_sampler(posterior, sample_shape: torch.Size, seed: int | None = None):
    return SobolQMCNormalSampler(sample_shape=sample_shape, seed=seed)
