"""FP64 restatement of BoTorch's GP-posterior / cached-root Sobol sampling / qLogNEI
path and of the exact marginal log-likelihood.

TEST INFRASTRUCTURE ONLY.  This module is the oracle that the HIP path is checked
against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the baseline legs of
``bench.py`` import it, and only as the checker or as the timed comparator - the
product package ``botorch_amd`` never imports it and has no code path through it.

Everything is plain device-agnostic PyTorch (float64 by default), written to follow
the reference's op sequence so that it doubles as the "plain FP64" comparator when
run on a GPU through PyTorch-ROCm ops.  Citations are to the reference tree
(botorch 0.18.1).  Functions marked [upstream] restate gpytorch / linear_operator
behaviour, which is not vendored in the reference; see each docstring for what
pins it in ``tests/test_oracle.py``.

Pinning summary (each function's check lives in tests/test_oracle.py):
  covar / sq_dist          closed form + brute-force direct differences; reference-held
                           Matern restatement models/fully_bayesian.py:99-126
  psd_safe_cholesky        torch.linalg.cholesky on PD input; jitter schedule on a
                           singular matrix (tiny eigen-problem solved in closed form)
  GP posterior             brute force with explicit inverse (n <= 16); LOO identity
                           cross_validation.py:644-679
  Sobol base samples       torch SobolEngine (same library as the reference), seeded
                           reproducibility test/sampling/test_normal.py:75-85
  sample_cached_cholesky   equals joint-Cholesky sampling (test/utils/test_low_rank.py)
  safe_math                closed forms of softplus / Cauchy / Pareto, properties from
                           test/utils/test_safe_math.py:260-467
  qLogNEI                  MockPosterior invariants test/acquisition/test_logei.py:339,
                           :404-405; cache_root == no-cache (:513-626); central FD
  MLL                      closed form -1/2(r'K^-1 r + log|K| + n log 2pi); adjoint
                           formula (Rasmussen & Williams eq. 5.9); central FD
"""
from __future__ import annotations

import math
import warnings
from dataclasses import dataclass, field

import torch
from torch import Tensor
from torch.nn.functional import softplus
from torch.quasirandom import SobolEngine

MATERN52 = 0
RBF = 1

TAU_RELU = 1e-6  # acquisition/logei.py:67
TAU_MAX = 1e-2  # acquisition/logei.py:68


class NotPSDError(RuntimeError):
    pass


class NanError(RuntimeError):
    pass


# ----------------------------------------------------------------------------------
# Kernels  (gpytorch MaternKernel(nu=2.5) / RBFKernel / ScaleKernel)  [upstream]
# ----------------------------------------------------------------------------------
def sq_dist(x1: Tensor, x2: Tensor) -> Tensor:
    """Squared distance the way gpytorch forms it [upstream]: centre by x1's mean,
    expand ||a||^2 + ||b||^2 - 2 a.b with one matmul, clamp at zero."""
    adjustment = x1.mean(-2, keepdim=True)
    x1 = x1 - adjustment
    x2 = x2 - adjustment
    x1_norm = x1.pow(2).sum(dim=-1, keepdim=True)
    x2_norm = x2.pow(2).sum(dim=-1, keepdim=True)
    x1_ = torch.cat([-2.0 * x1, x1_norm, torch.ones_like(x1_norm)], dim=-1)
    x2_ = torch.cat([x2, torch.ones_like(x2_norm), x2_norm], dim=-1)
    res = x1_.matmul(x2_.transpose(-2, -1))
    return res.clamp_min(0)


def covar(kind: int, x1: Tensor, x2: Tensor, lengthscale: Tensor,
          outputscale: Tensor | float = 1.0) -> Tensor:
    """k(x1, x2) for an ARD Matern-5/2 or RBF kernel times ``outputscale``.

    Matern-5/2 follows gpytorch's MaternKernel.forward [upstream] and the
    reference-held restatement models/fully_bayesian.py:99-103,115-126:
    x <- (x - mean(x1)) / l, r = sqrt(clamp(sq_dist, 1e-30)),
    k = (1 + sqrt5 r + 5/3 r^2) exp(-sqrt5 r).  RBF: exp(-r^2/2).
    """
    mean = x1.reshape(-1, x1.size(-1)).mean(0)
    x1_ = (x1 - mean) / lengthscale
    x2_ = (x2 - mean) / lengthscale
    d2 = sq_dist(x1_, x2_)
    if kind == MATERN52:
        dist = d2.clamp_min(1e-30).sqrt()
        k = (math.sqrt(5) * dist).add(1).add(5.0 / 3.0 * dist**2) * torch.exp(
            -math.sqrt(5) * dist)
    elif kind == RBF:
        k = torch.exp(-0.5 * d2)
    else:
        raise ValueError(f"unknown kernel kind {kind}")
    return k * outputscale


def covar_direct(kind: int, x1: Tensor, x2: Tensor, lengthscale: Tensor,
                 outputscale: float = 1.0) -> Tensor:
    """Brute-force kernel from direct coordinate differences (pins ``covar``)."""
    diff = (x1.unsqueeze(-2) - x2.unsqueeze(-3)) / lengthscale
    d2 = (diff**2).sum(-1)
    r = d2.clamp_min(1e-30).sqrt()
    if kind == MATERN52:
        k = (1 + math.sqrt(5) * r + 5.0 / 3.0 * r**2) * torch.exp(-math.sqrt(5) * r)
    else:
        k = torch.exp(-0.5 * d2)
    return k * outputscale


# ----------------------------------------------------------------------------------
# psd_safe_cholesky  (linear_operator, called at utils/low_rank.py:140,
# acquisition/cached_cholesky.py:123; cholesky_max_tries=6 at botorch/__init__.py:54)
# ----------------------------------------------------------------------------------
def psd_safe_cholesky(A: Tensor, max_tries: int = 6, jitter: float | None = None,
                      warn: bool = True) -> Tensor:
    """Cholesky with jitter escalation [upstream]: cholesky_ex; if any batch entry
    fails, NaN check, then add jitter*10^i increments to the failing entries only,
    up to ``max_tries`` times, else NotPSDError."""
    L, info = torch.linalg.cholesky_ex(A)
    if not torch.any(info):
        return L
    if torch.isnan(A).any():
        raise NanError("cholesky_cpu: input matrix contains NaNs")
    if jitter is None:
        jitter = 1e-6 if A.dtype == torch.float32 else 1e-8
    Aprime = A.clone()
    jitter_prev = 0.0
    for i in range(max_tries):
        jitter_new = jitter * (10**i)
        diag_add = ((info > 0) * (jitter_new - jitter_prev)).unsqueeze(-1).expand(
            *Aprime.shape[:-1])
        Aprime.diagonal(dim1=-1, dim2=-2).add_(diag_add)
        jitter_prev = jitter_new
        if warn:
            warnings.warn(f"A not p.d., added jitter of {jitter_new:.1e} to the diagonal",
                          RuntimeWarning)
        L, info = torch.linalg.cholesky_ex(Aprime)
        if not torch.any(info):
            return L
    raise NotPSDError(f"Matrix not positive definite after repeatedly adding jitter up "
                      f"to {jitter_new:.1e}.")


# ----------------------------------------------------------------------------------
# Standardize  (models/transforms/outcome.py:312-360 forward, :431-511 untransform)
# ----------------------------------------------------------------------------------
def standardize(Y: Tensor, min_stdv: float = 1e-8):
    """Returns (Y_tf, means, stdvs) for an n x 1 tensor, outcome.py:332-352."""
    if Y.shape[-2] == 1:
        stdvs = torch.ones_like(Y[..., :1, :])
    else:
        stdvs = Y.std(dim=-2, keepdim=True)
    stdvs = stdvs.where(stdvs >= min_stdv, torch.full_like(stdvs, 1.0))
    means = Y.mean(dim=-2, keepdim=True)
    return (Y - means) / stdvs, means, stdvs


# ----------------------------------------------------------------------------------
# SingleTaskGP exact prediction with fast_pred_var (models/gpytorch.py:544-610,
# models/utils/assorted.py:305-315, gpytorch DefaultPredictionStrategy [upstream])
# ----------------------------------------------------------------------------------
@dataclass
class OracleGP:
    train_X: Tensor
    train_Y: Tensor  # n x 1, original scale
    kind: int = MATERN52
    lengthscale: Tensor | None = None
    outputscale: float = 1.0
    noise: float = 1e-4
    constant: float = 0.0
    standardize_Y: bool = True
    # caches
    y_mean: Tensor = field(init=False)
    y_std: Tensor = field(init=False)
    train_targets: Tensor = field(init=False)
    L: Tensor = field(init=False)
    mean_cache: Tensor = field(init=False)
    covar_cache: Tensor = field(init=False)

    def __post_init__(self):
        if self.lengthscale is None:
            self.lengthscale = torch.ones(self.train_X.shape[-1], dtype=self.train_X.dtype,
                                          device=self.train_X.device)
        if self.standardize_Y:
            Yt, m, s = standardize(self.train_Y)
        else:
            Yt = self.train_Y
            m = torch.zeros_like(self.train_Y[..., :1, :])
            s = torch.ones_like(m)
        self.y_mean, self.y_std = m.reshape(()), s.reshape(())
        self.train_targets = Yt.squeeze(-1)
        self.refresh()

    def refresh(self):
        """Caches mean_cache = K^-1 (y - c) and covar_cache R = L^-T (fast_pred_var)."""
        K = self.train_covar()
        self.L = psd_safe_cholesky(K)
        r = (self.train_targets - self.constant).unsqueeze(-1)
        self.mean_cache = torch.cholesky_solve(r, self.L).squeeze(-1)
        eye = torch.eye(K.shape[-1], dtype=K.dtype, device=K.device)
        Linv = torch.linalg.solve_triangular(self.L, eye, upper=False)
        self.covar_cache = Linv.transpose(-1, -2)

    def train_covar(self) -> Tensor:
        K = covar(self.kind, self.train_X, self.train_X, self.lengthscale, self.outputscale)
        n = K.shape[-1]
        return K + self.noise * torch.eye(n, dtype=K.dtype, device=K.device)

    def posterior_std(self, X: Tensor):
        """Mean (... x m) and covariance (... x m x m) in the standardized space."""
        Kxt = covar(self.kind, X, self.train_X, self.lengthscale, self.outputscale)
        Kxx = covar(self.kind, X, X, self.lengthscale, self.outputscale)
        mean = self.constant + Kxt.matmul(self.mean_cache)
        root = Kxt.matmul(self.covar_cache)
        cov = Kxx - root.matmul(root.transpose(-1, -2))
        return mean, cov

    def posterior(self, X: Tensor, observation_noise=False):
        """Untransformed posterior (mean ... x m, covariance ... x m x m)
        following Standardize.untransform_posterior (outcome.py:488-507).
        observation_noise (models/gpytorch.py:514-538, applied before the
        untransform): True adds the likelihood's noise to the diagonal, a tensor
        (... x m x 1) adds its own values."""
        mean, cov = self.posterior_std(X)
        if observation_noise is not False:
            m = X.shape[-2]
            if torch.is_tensor(observation_noise):
                nz = observation_noise.squeeze(-1).to(cov).expand(*cov.shape[:-1])
            else:
                nz = torch.full(cov.shape[:-1], float(self.noise), dtype=cov.dtype,
                                device=cov.device)
            cov = cov + torch.diag_embed(nz)
            assert cov.shape[-1] == m
        return self.y_mean + self.y_std * mean, (self.y_std**2) * cov


# ----------------------------------------------------------------------------------
# Sobol QMC base samples (sampling/qmc.py:59-97, utils/sampling.py:114-143,
# sampling/normal.py:68-135,173-213)
# ----------------------------------------------------------------------------------
def draw_sobol_normal_samples(d: int, n: int, seed: int, dtype=torch.float64) -> Tensor:
    """Owen-scrambled Sobol + inverse CDF with the eps clamp of qmc.py:80-83."""
    eng = SobolEngine(dimension=d, scramble=True, seed=seed)
    u = eng.draw(n, dtype=dtype)
    v = 0.5 + (1 - torch.finfo(u.dtype).eps) * (u - 0.5)
    return torch.erfinv(2 * v - 1) * math.sqrt(2)


def cached_root_base_samples(S: int, n_b: int, q: int, seed: int,
                             dtype=torch.float64) -> Tensor:
    """Base samples for the joint (X_baseline, X) posterior with the baseline
    columns taken from the baseline sampler (sampling/normal.py:128-131):
    Z = [Sobol_{n_b}(seed) | Sobol_{n_b+q}(seed)[:, n_b:]]  (S x (n_b+q))."""
    Z = draw_sobol_normal_samples(n_b + q, S, seed, dtype)
    Z[:, :n_b] = draw_sobol_normal_samples(n_b, S, seed, dtype)
    return Z


# ----------------------------------------------------------------------------------
# safe_math  (utils/safe_math.py)
# ----------------------------------------------------------------------------------
def cauchy(x: Tensor) -> Tensor:  # safe_math.py:461-463
    return 1 / (1 + x.square())


def fatplus(x: Tensor, tau: float) -> Tensor:  # safe_math.py:304-325
    z = x / tau
    return tau * (softplus(z) + 0.1 * cauchy(z))


def log_fatplus(x: Tensor, tau: float) -> Tensor:  # safe_math.py:298-301
    return fatplus(x, tau).log()


def pareto(x: Tensor, alpha: float = 2.0) -> Tensor:  # safe_math.py:466-490
    alpha = alpha / 2
    beta_1 = 2 * alpha
    beta_0 = alpha * beta_1
    return (beta_0 / (beta_0 + beta_1 * x + x.square())).pow(alpha)


def _inf_max_helper(max_fun, x: Tensor, dim: int) -> Tensor:  # safe_math.py:146-191
    if x.shape[dim] == 0:
        return x.sum(dim=dim) - torch.inf
    M = x.amax(dim=dim, keepdim=True)
    is_inf_max = torch.logical_and(*torch.broadcast_tensors(M.isinf(), x == M))
    has_inf_max = is_inf_max.any(dim=dim, keepdim=True)
    y_inf = x.masked_fill(~is_inf_max, 0.0)
    M_no_inf = M.masked_fill(M.isinf(), 0.0)
    y_no_inf = x.masked_fill(has_inf_max, 0.0) - M_no_inf
    res = torch.where(has_inf_max, y_inf.sum(dim=dim, keepdim=True),
                      M_no_inf + max_fun(y_no_inf, dim))
    return res.squeeze(dim)


def fatmax(x: Tensor, dim: int, tau: float, alpha: float = 2.0) -> Tensor:
    """safe_math.py:328-355."""
    def max_fun(y, dim):
        return tau * pareto(-y / tau, alpha=alpha).sum(dim=dim, keepdim=True).log()
    return _inf_max_helper(max_fun, x, dim)


def logsumexp(x: Tensor, dim: int) -> Tensor:  # safe_math.py:123-143
    return _inf_max_helper(lambda y, dim: torch.logsumexp(y, dim=dim, keepdim=True), x, dim)


def logmeanexp(x: Tensor, dim: int) -> Tensor:  # safe_math.py:213-225
    return logsumexp(x, dim) - math.log(x.shape[dim])


# ----------------------------------------------------------------------------------
# sample_cached_cholesky  (utils/low_rank.py:84-172)
# ----------------------------------------------------------------------------------
def sample_cached_cholesky(mean_q: Tensor, bottom_rows: Tensor, baseline_L: Tensor,
                           Z: Tensor, max_tries: int = 6) -> Tensor:
    """mean_q: b x q, bottom_rows: b x q x (n_b+q) rows of the joint covariance,
    baseline_L: n_b x n_b, Z: S x (n_b+q).  Returns samples S x b x q."""
    q = mean_q.shape[-1]
    bl, br = bottom_rows.split([bottom_rows.shape[-1] - q, q], dim=-1)
    bl_chol = torch.linalg.solve_triangular(baseline_L, bl.transpose(-2, -1),
                                            upper=False).transpose(-2, -1)
    br_to_chol = br - bl_chol @ bl_chol.transpose(-2, -1)
    br_chol = psd_safe_cholesky(br_to_chol, max_tries=max_tries)
    new_Lq = torch.cat([bl_chol, br_chol], dim=-1)  # b x q x m
    res = torch.einsum("bqm,sm->sbq", new_Lq, Z) + mean_q.unsqueeze(0)
    if torch.isnan(res).any() or torch.isinf(res).any():
        raise NanError("Samples contain nans or infs.")
    return res


# ----------------------------------------------------------------------------------
# qLogNoisyExpectedImprovement  (acquisition/logei.py:246-585, monte_carlo.py:268-305,
# cached_cholesky.py:98-192)
# ----------------------------------------------------------------------------------
class OracleQLogNEI:
    """qLogNEI with cache_root=True, incremental, fat=True, IdentityMCObjective,
    no constraints and no pending points, with an already-pruned X_baseline
    (prune_baseline=False)."""

    def __init__(self, gp: OracleGP, X_baseline: Tensor, num_samples: int = 512,
                 seed: int = 0, tau_relu: float = TAU_RELU, tau_max: float = TAU_MAX,
                 cache_root: bool = True):
        self.cache_root = cache_root
        self.gp = gp
        self.X_baseline = X_baseline
        self.S = num_samples
        self.seed = seed
        self.tau_relu = tau_relu
        self.tau_max = tau_max
        n_b = X_baseline.shape[-2]
        with torch.no_grad():  # logei.py:431-447
            mean_b, cov_b = gp.posterior(X_baseline)
            self.baseline_L = psd_safe_cholesky(cov_b)  # cached_cholesky.py:98-124
            Zb = draw_sobol_normal_samples(n_b, num_samples, seed, X_baseline.dtype).to(
                X_baseline.device)
            self.baseline_samples = mean_b.unsqueeze(0) + Zb @ self.baseline_L.T
            self.baseline_best_f = self.baseline_samples.amax(dim=-1)  # acquisition/utils.py:134-138
        self._Z_cache = {}

    def base_samples(self, q: int) -> Tensor:
        if q not in self._Z_cache:
            n_b = self.X_baseline.shape[-2]
            self._Z_cache[q] = cached_root_base_samples(
                self.S, n_b, q, self.seed, self.X_baseline.dtype).to(self.X_baseline.device)
        return self._Z_cache[q]

    def samples(self, X: Tensor) -> Tensor:
        """S x b x q posterior samples at X (b x q x d) via the reference's
        sequence: joint posterior over X_full = [X_baseline, X]  (logei.py:527-565)."""
        b, q, _ = X.shape
        X_full = torch.cat([self.X_baseline.expand(b, -1, -1), X], dim=-2)
        mean, cov = self.gp.posterior(X_full)
        if not self.cache_root:
            # the joint path (cached_cholesky.py:160-171 -> sampler(posterior)):
            # mean + L z over all n_b + q points with L = psd_safe_cholesky of the
            # joint covariance, then the last q columns
            L = psd_safe_cholesky(cov)
            full = torch.einsum("bij,sj->sbi", L, self.base_samples(q)) + mean.unsqueeze(0)
            return full[..., -q:]
        bottom_rows = cov[..., -q:, :]
        return sample_cached_cholesky(mean[..., -q:], bottom_rows, self.baseline_L,
                                      self.base_samples(q))

    def reduce(self, Y: Tensor) -> Tensor:
        """Y: S x b x q -> b   (logei.py:376-391,688-715; monte_carlo.py:268-290)."""
        best_f = self.baseline_best_f.view(-1, 1, 1).to(Y)
        v = log_fatplus(Y - best_f, tau=self.tau_relu)
        u = fatmax(v, dim=-1, tau=self.tau_max)
        return logmeanexp(u, dim=0)

    def __call__(self, X: Tensor) -> Tensor:
        if X.dim() == 2:
            X = X.unsqueeze(0)
        return self.reduce(self.samples(X))

    def value_and_grad(self, X: Tensor):
        """One optimiser step's worth of work (generation/gen.py:466-469)."""
        X = X.detach().clone().requires_grad_(True)
        val = self(X)
        (g,) = torch.autograd.grad(val.sum(), X)
        return val.detach(), g


# ----------------------------------------------------------------------------------
# prune_inferior_points (acquisition/utils.py:347-437) with a fixed-seed sampler
# ----------------------------------------------------------------------------------
def prune_inferior_points(gp: OracleGP, X: Tensor, num_samples: int = 2048,
                          seed: int = 0, max_frac: float = 1.0) -> Tensor:
    max_points = math.ceil(max_frac * X.size(-2))
    with torch.no_grad():
        mean, cov = gp.posterior(X)
        L = psd_safe_cholesky(cov)
        Z = draw_sobol_normal_samples(X.shape[-2], num_samples, seed, X.dtype).to(X.device)
        samples = mean.unsqueeze(0) + Z @ L.T  # S x n
    is_best = torch.argmax(samples, dim=-1)
    idcs, counts = torch.unique(is_best, return_counts=True)
    if len(idcs) > max_points:
        counts, order_idcs = torch.sort(counts, stable=True, descending=True)
        idcs = order_idcs[:max_points]
    return X[idcs], idcs


# ----------------------------------------------------------------------------------
# Exact marginal log likelihood (gpytorch ExactMarginalLogLikelihood [upstream],
# optim/closures/model_closures.py:125-145)
# ----------------------------------------------------------------------------------
def lognormal_log_prob(x: Tensor, loc: float, scale: float) -> Tensor:
    """log density of LogNormal(loc, scale) at x (torch.distributions.LogNormal)."""
    lx = torch.log(x)
    return (-((lx - loc) ** 2) / (2 * scale**2) - math.log(scale)
            - 0.5 * math.log(2 * math.pi) - lx)


def exact_mll(kind: int, X: Tensor, y: Tensor, lengthscale: Tensor, outputscale: Tensor,
              noise: Tensor, constant: Tensor, priors: dict | None = None) -> Tensor:
    """(log N(y; c, K + noise I) + sum log p(theta)) / n, computed with
    psd_safe_cholesky like the reference (forced Cholesky, botorch/__init__.py:51-56)."""
    n = X.shape[-2]
    K = covar(kind, X, X, lengthscale, outputscale)
    Khat = K + noise * torch.eye(n, dtype=X.dtype, device=X.device)
    L = psd_safe_cholesky(Khat)
    r = (y - constant).unsqueeze(-1)
    a = torch.linalg.solve_triangular(L, r, upper=False)
    quad = (a**2).sum()
    logdet = 2 * torch.log(torch.diagonal(L)).sum()
    ll = -0.5 * (quad + logdet + n * math.log(2 * math.pi))
    if priors:
        if "lengthscale" in priors:
            loc, scale = priors["lengthscale"]
            ll = ll + lognormal_log_prob(lengthscale, loc, scale).sum()
        if "noise" in priors:
            loc, scale = priors["noise"]
            ll = ll + lognormal_log_prob(noise.reshape(1), loc, scale).sum()
        if "outputscale" in priors:
            loc, scale = priors["outputscale"]
            ll = ll + lognormal_log_prob(outputscale.reshape(1), loc, scale).sum()
    return ll / n


def exact_mll_adjoint(kind: int, X: Tensor, y: Tensor, lengthscale: Tensor,
                      outputscale: float, noise: float, constant: float):
    """Closed-form gradient of the (un-normalised, prior-free) log likelihood
    w.r.t. (lengthscale[d], outputscale, noise, constant):
    dL/dtheta = 1/2 tr((a a^T - K^-1) dK/dtheta)  (Rasmussen & Williams eq. 5.9)."""
    n = X.shape[-2]
    with torch.no_grad():
        K = covar(kind, X, X, lengthscale, outputscale)
        Khat = K + noise * torch.eye(n, dtype=X.dtype, device=X.device)
        L = torch.linalg.cholesky(Khat)
        a = torch.cholesky_solve((y - constant).unsqueeze(-1), L).squeeze(-1)
        Kinv = torch.cholesky_inverse(L)
        W = torch.outer(a, a) - Kinv
        diff = X.unsqueeze(-2) - X.unsqueeze(-3)  # n x n x d
        r2 = ((diff / lengthscale) ** 2).sum(-1)
        r = r2.sqrt()
        if kind == MATERN52:
            h = outputscale * (5.0 / 3.0) * (1 + math.sqrt(5) * r) * torch.exp(-math.sqrt(5) * r)
        else:
            h = outputscale * torch.exp(-0.5 * r2)
        g_ls = 0.5 * torch.einsum("ij,ij,ijd->d", W, h, diff**2) / lengthscale**3
        g_os = 0.5 * (W * K).sum() / outputscale
        g_noise = 0.5 * torch.diagonal(W).sum()
        g_c = a.sum()
    return g_ls, g_os, g_noise, g_c


# ----------------------------------------------------------------------------------
# qLogEHVI log-areas (acquisition/multi_objective/logei.py:437-479, steps 1-4 of
# _compute_log_qehvi, the computation csrc/logei_fused.cpp fuses)
# ----------------------------------------------------------------------------------
def fatmin(x: Tensor, dim: int, tau: float, alpha: float = 2.0) -> Tensor:
    """safe_math.py:358-381: -fatmax(-x)."""
    return -fatmax(-x, dim=dim, tau=tau, alpha=alpha)


def log_areas(obj_subsets: Tensor, cell_lower: Tensor, cell_upper: Tensor,
              tau_relu: float = TAU_RELU, tau_max: float = TAU_MAX) -> Tensor:
    """obj_subsets: B x n_sub x i x m; cell bounds nc x m or B x nc x m.
    Returns B x nc x n_sub: sum_k fatmin(fatmin_j(log_fatplus(obj - l_c)), log(u_c - l_c))
    with u clamped at 1e10 (FP64) / 1e8 (FP32), as _log_improvement, _smooth_min,
    _log_cell_lengths and the sum over objectives compute it."""
    cl = cell_lower if cell_lower.dim() == 3 else cell_lower.unsqueeze(0)
    cu = cell_upper if cell_upper.dim() == 3 else cell_upper.unsqueeze(0)
    Z = obj_subsets.unsqueeze(1) - cl[:, :, None, None, :]     # B x nc x n_sub x i x m
    li = log_fatplus(Z, tau=tau_relu)
    lim = fatmin(li, dim=-2, tau=tau_max)                       # B x nc x n_sub x m
    clamp = 1e10 if obj_subsets.dtype == torch.float64 else 1e8
    lcl = (cu.clamp_max(clamp) - cl).log()[:, :, None, :]       # B' x nc x 1 x m
    lm = fatmin(torch.stack(torch.broadcast_tensors(lim, lcl), dim=-1), dim=-1, tau=tau_max)
    return lm.sum(dim=-1)
