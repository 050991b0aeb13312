"""Exact marginal log likelihood and fit_gpytorch_mll on libbotorch_amd.

Mirrors gpytorch's ExactMarginalLogLikelihood as used by BoTorch
(optim/closures/model_closures.py:125-145: loss = -(log N(y; c, K) + sum of
hyperparameter log-priors) / n on the standardized targets) and fit.py:76-289 /
optim/fit.py:46-141 (L-BFGS-B over the raw parameters with the constraint
bounds of optim/utils/model_utils.py:68-108, retries on numerical failure).

Value and gradient of the log likelihood come from one library call
(bta_mll_value_grad: Cholesky, triangular inverse and the fused
1/2 tr((a a^T - K^-1) dK/dtheta) kernel); priors and the chain rule through the
parameter constraints are O(d) host arithmetic.
"""
from __future__ import annotations

import ctypes
import math
import warnings

import numpy as np
import scipy.optimize
import torch

from . import _lib
from .exceptions import NanError, NotPSDError
from .gp import ScaleKernel, SingleTaskGP


class ExactMarginalLogLikelihood(torch.nn.Module):
    def __init__(self, likelihood, model: SingleTaskGP):
        super().__init__()
        self.likelihood = likelihood
        self.model = model
        self._ws = None

    # -- parameters -----------------------------------------------------------
    def _params(self):
        """(name, module, attribute) of every trainable raw parameter."""
        m = self.model
        k = m.covar_module
        base = k.base_kernel if isinstance(k, ScaleKernel) else k
        out = [("lengthscale", base, "lengthscale")]
        if isinstance(k, ScaleKernel):
            out.append(("outputscale", k, "outputscale"))
        out.append(("noise", m.likelihood, "noise"))
        out.append(("constant", m.mean_module, "constant"))
        return out

    def raw_parameters(self):
        return [(n, getattr(mod, f"raw_{attr}")) for n, mod, attr in self._params()]

    # -- value and gradient -----------------------------------------------------
    def loglik_and_grad(self):
        """Un-normalised log N(y; c, K) and its gradient w.r.t. the natural
        parameters (lengthscale d, outputscale, noise, constant)."""
        L = _lib.lib()
        m = self.model
        X = m.train_inputs[0].contiguous()
        y = m.train_targets.contiguous()
        n, d = X.shape
        hp = m.hyperparameters()
        ls = hp["lengthscale"].cpu().tolist()
        need = L.bta_mll_workspace_size(n, d)
        if self._ws is None or self._ws.numel() < need:
            self._ws = _lib.workspace(need, X.device)
        out = (ctypes.c_double * (4 + d))()
        lsh = (ctypes.c_double * d)(*ls)
        _lib.check(L.bta_mll_value_grad(m.kernel_kind, _lib.ptr(X), _lib.ptr(y), n, d, lsh,
                                        hp["outputscale"], hp["noise"], hp["constant"], 6, out,
                                        _lib.ptr(self._ws), self._ws.numel(), _lib.stream()),
                   "mll_value_grad")
        o = list(out)
        return o[0], {"outputscale": o[1], "noise": o[2], "constant": o[3],
                      "lengthscale": np.array(o[4:4 + d])}

    def loss_and_grads(self):
        """loss = -(loglik + sum log priors) / n and d loss / d raw parameter."""
        ll, g = self.loglik_and_grad()
        n = self.model.train_inputs[0].shape[0]
        total = ll
        grads = []
        for name, mod, attr in self._params():
            raw = getattr(mod, f"raw_{attr}").detach().cpu()
            if name == "constant":
                dv = torch.tensor(g["constant"], dtype=torch.float64)
                grads.append(-dv / n)
                continue
            c = getattr(mod, f"raw_{attr}_constraint")
            val = c.transform_value(raw)
            dv = torch.as_tensor(g[name], dtype=torch.float64).reshape(raw.shape)
            prior = getattr(mod, f"{attr}_prior", None)
            if prior is not None:
                total += float(prior.log_prob(val).sum())
                dv = dv + prior.dlog_prob(val)
            grads.append(-(dv * c.dvalue_draw(raw)) / n)
        return -total / n, grads

    def forward(self, *args, **kwargs):
        loss, _ = self.loss_and_grads()
        return torch.tensor(-loss, dtype=torch.float64)


def get_parameters_and_bounds(mll: ExactMarginalLogLikelihood):
    """Flattened raw parameters and their L-BFGS-B bounds
    (optim/utils/model_utils.py:68-108)."""
    xs, lo, hi = [], [], []
    for name, mod, attr in mll._params():
        raw = getattr(mod, f"raw_{attr}").detach().cpu().reshape(-1)
        xs.append(raw.numpy())
        c = getattr(mod, f"raw_{attr}_constraint", None)
        lb, ub = c.raw_bounds() if c is not None else (None, None)
        lo.extend([-np.inf if lb is None else lb] * raw.numel())
        hi.extend([np.inf if ub is None else ub] * raw.numel())
    return np.concatenate(xs).astype(np.float64), np.array(lo), np.array(hi)


def _set_flat(mll: ExactMarginalLogLikelihood, x: np.ndarray) -> None:
    i = 0
    for name, mod, attr in mll._params():
        raw = getattr(mod, f"raw_{attr}")
        k = raw.numel()
        with torch.no_grad():
            raw.copy_(torch.from_numpy(x[i:i + k]).reshape(raw.shape).to(raw))
        i += k


def fit_gpytorch_mll_scipy(mll: ExactMarginalLogLikelihood, options: dict | None = None):
    """One L-BFGS-B run over the raw parameters (optim/fit.py:46-141)."""
    options = {"maxiter": 200, **(options or {})}
    x0, lo, hi = get_parameters_and_bounds(mll)
    nfev = [0]

    def fun(x):
        _set_flat(mll, x)
        nfev[0] += 1
        try:
            loss, grads = mll.loss_and_grads()
        except (NotPSDError, NanError):
            return float("nan"), np.full_like(x, float("nan"))
        return loss, np.concatenate([g.reshape(-1).numpy() for g in grads])

    bounds = list(zip(np.where(np.isfinite(lo), lo, None), np.where(np.isfinite(hi), hi, None)))
    res = scipy.optimize.minimize(fun, x0, jac=True, method="L-BFGS-B", bounds=bounds,
                                  options=options)
    _set_flat(mll, res.x)
    return res


def _sample_all_priors(model: SingleTaskGP, generator: torch.Generator) -> None:
    """Re-initialise every hyperparameter that has a prior with a prior draw."""
    for mod in model.modules():
        for attr in ("lengthscale", "outputscale", "noise"):
            prior = getattr(mod, f"{attr}_prior", None)
            raw = getattr(mod, f"raw_{attr}", None)
            if prior is None or raw is None:
                continue
            c = getattr(mod, f"raw_{attr}_constraint")
            v = prior.sample(raw.shape, generator=generator)
            if c.transform is None:
                v = v.clamp_min(c.lower_bound)
            with torch.no_grad():
                raw.copy_(c.inverse_transform(v.to(raw)).to(raw))


def fit_gpytorch_mll(mll: ExactMarginalLogLikelihood, optimizer_kwargs: dict | None = None,
                     max_attempts: int = 5, seed: int = 0, **kwargs):
    """fit.py:76-289 (_fit_fallback): retry from prior draws on numerical
    failure or a non-finite result, keep the best attempt."""
    model = mll.model
    state0 = {k: v.detach().clone() for k, v in model.state_dict().items()}
    gen = torch.Generator().manual_seed(seed)
    best = None
    for attempt in range(1, max_attempts + 1):
        if attempt > 1:
            model.load_state_dict(state0)
            _sample_all_priors(model, gen)
        try:
            res = fit_gpytorch_mll_scipy(mll, (optimizer_kwargs or {}).get("options"))
        except (NotPSDError, NanError) as e:
            warnings.warn(f"Fit attempt #{attempt} failed: {e}", RuntimeWarning)
            continue
        if not np.isfinite(res.fun):
            warnings.warn(f"Fit attempt #{attempt} produced a non-finite loss", RuntimeWarning)
            continue
        if best is None or res.fun < best[0]:
            best = (res.fun, {k: v.detach().clone() for k, v in model.state_dict().items()})
        if res.success:
            break
    if best is None:
        model.load_state_dict(state0)
        raise RuntimeError("All attempts to fit the model have failed.")
    model.load_state_dict(best[1])
    model.eval()
    return mll
