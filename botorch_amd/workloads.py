"""Seeded synthetic problems of the benchmark configurations.

Each generator returns CPU float64 tensors from its own torch.Generator, so
the same inputs feed the HIP path, the FP64 oracle and the fixtures.

  gp_small   : n=64, d=6, X ~ U[0,1]^6 (seed 0), y = sum sin(3x) + 0.1|x-0.5|^2 + N(0, 0.01^2),
               Matern-5/2, lengthscale 0.5, outputscale 1, noise 1e-4;
               qLogNEI q=1, 512 Sobol samples, 32 restarts.
  gp_headline: n=2048, d=12, X ~ U[0,1]^12, y = sum sin(3x) + N(0, 0.01^2),
               lengthscales ~ U[0.2, 1], outputscale 1, noise 1e-4;
               qLogNEI q=4, 512 restarts, 512 samples.
  gp_mll     : n=4096, d=16, y = GP prior draw (lengthscale 0.4) + N(0, 0.05^2).
  gp_large   : n=16384, d=20, X ~ U[0,1]^20 with 10% of the points N(c, 0.02^2) around
               one centre c ~ U[0.3,0.7]^20, y = sum sin(3x) + N(0, 0.01^2),
               lengthscales ~ U[0.2, 1], noise 1e-4 (the likelihood's lower bound);
               qLogNEI q=8, 1024 restarts, 256 samples.
  gp_batch   : 64 independent models, n log-uniform in [128, 4096] (multiples of 8),
               d uniform in {4..20}, each 256 restarts x q=2 with 256 samples.
"""
from __future__ import annotations

import math

import torch


def gp_small(seed: int = 0):
    g = torch.Generator().manual_seed(seed)
    n, d = 64, 6
    X = torch.rand(n, d, generator=g, dtype=torch.float64)
    y = (torch.sin(3 * X).sum(-1) + 0.1 * ((X - 0.5) ** 2).sum(-1)
         + 0.01 * torch.randn(n, generator=g, dtype=torch.float64))
    Xq = torch.rand(32, 1, d, generator=g, dtype=torch.float64)
    return {"X": X, "Y": y.unsqueeze(-1), "lengthscale": torch.full((d,), 0.5, dtype=torch.float64),
            "outputscale": 1.0, "noise": 1e-4, "Xq": Xq, "q": 1, "S": 512, "b": 32}


def gp_headline(seed: int = 0, n: int = 2048, d: int = 12, b: int = 512, q: int = 4,
                S: int = 512):
    g = torch.Generator().manual_seed(seed)
    X = torch.rand(n, d, generator=g, dtype=torch.float64)
    y = torch.sin(3 * X).sum(-1) + 0.01 * torch.randn(n, generator=g, dtype=torch.float64)
    ls = 0.2 + 0.8 * torch.rand(d, generator=g, dtype=torch.float64)
    Xq = torch.rand(b, q, d, generator=g, dtype=torch.float64)
    return {"X": X, "Y": y.unsqueeze(-1), "lengthscale": ls, "outputscale": 1.0,
            "noise": 1e-4, "Xq": Xq, "q": q, "S": S, "b": b}


def gp_mll(seed: int = 0, n: int = 4096, d: int = 16, lengthscale: float = 0.4,
           noise_sd: float = 0.05):
    """Training data whose y is a draw from a Matern-5/2 GP prior plus noise."""
    g = torch.Generator().manual_seed(seed)
    X = torch.rand(n, d, generator=g, dtype=torch.float64)
    z = torch.randn(n, generator=g, dtype=torch.float64)
    eps = torch.randn(n, generator=g, dtype=torch.float64)
    return {"X": X, "z": z, "eps": eps, "lengthscale": lengthscale, "noise_sd": noise_sd}


def prior_draw(X: torch.Tensor, z: torch.Tensor, lengthscale: float) -> torch.Tensor:
    """f = chol(K + 1e-6 I) z for a Matern-5/2 kernel (any device)."""
    diff = (X.unsqueeze(-2) - X.unsqueeze(-3)) / lengthscale
    r = diff.pow(2).sum(-1).clamp_min(1e-30).sqrt()
    K = (1 + math.sqrt(5) * r + 5.0 / 3.0 * r**2) * torch.exp(-math.sqrt(5) * r)
    K = K + 1e-6 * torch.eye(X.shape[0], dtype=X.dtype, device=X.device)
    return torch.linalg.cholesky(K) @ z


def gp_large(seed: int = 0, n: int = 16384, d: int = 20, b: int = 1024, q: int = 8,
             S: int = 256, cluster_frac: float = 0.1):
    g = torch.Generator().manual_seed(seed)
    X = torch.rand(n, d, generator=g, dtype=torch.float64)
    nc = int(cluster_frac * n)
    c = 0.3 + 0.4 * torch.rand(d, generator=g, dtype=torch.float64)
    X[:nc] = c + 0.02 * torch.randn(nc, d, generator=g, dtype=torch.float64)
    y = torch.sin(3 * X).sum(-1) + 0.01 * torch.randn(n, generator=g, dtype=torch.float64)
    ls = 0.2 + 0.8 * torch.rand(d, generator=g, dtype=torch.float64)
    Xq = torch.rand(b, q, d, generator=g, dtype=torch.float64)
    return {"X": X, "Y": y.unsqueeze(-1), "lengthscale": ls, "outputscale": 1.0,
            "noise": 1e-4, "Xq": Xq, "q": q, "S": S, "b": b}


def gp_batch(seed: int = 0, count: int = 64, b: int = 256, q: int = 2, S: int = 256):
    g = torch.Generator().manual_seed(seed)
    out = []
    for _ in range(count):
        u = torch.rand((), generator=g, dtype=torch.float64).item()
        n = int(round(math.exp(math.log(128) + u * (math.log(4096) - math.log(128))) / 8)) * 8
        d = int(torch.randint(4, 21, (), generator=g).item())
        X = torch.rand(n, d, generator=g, dtype=torch.float64)
        y = torch.sin(3 * X).sum(-1) + 0.01 * torch.randn(n, generator=g, dtype=torch.float64)
        ls = 0.2 + 0.8 * torch.rand(d, generator=g, dtype=torch.float64)
        Xq = torch.rand(b, q, d, generator=g, dtype=torch.float64)
        out.append({"X": X, "Y": y.unsqueeze(-1), "lengthscale": ls, "outputscale": 1.0,
                    "noise": 1e-4, "Xq": Xq, "q": q, "S": S, "b": b})
    return out
