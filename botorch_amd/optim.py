"""optimize_acqf: multi-start L-BFGS-B over the batched qLogNEI value+grad.

Mirrors optim/optimize.py:619-845 (optimize_acqf -> _optimize_acqf_batch),
optim/initializers.py:258-483,959 (gen_batch_initial_conditions,
initialize_q_batch), generation/gen.py:62-507,680-709 (gen_candidates_scipy,
f_np_wrapper, get_best_candidates) and optim/batched_lbfgs_b.py:169-636 (lock-step
batched L-BFGS-B: one SciPy L-BFGS-B state per restart, one batched
function+gradient evaluation per step for all restarts that asked for one).

Each optimiser step issues ONE library call for all active restarts
(``qLogNoisyExpectedImprovement.value_and_grad``); the per-restart L-BFGS-B
bookkeeping runs on the host in SciPy's own L-BFGS-B routine, as in the
reference.
"""
from __future__ import annotations

import math
import time
import warnings

import numpy as np
import torch
from scipy.optimize import _lbfgsb
from torch import Tensor

from .exceptions import OptimizationGradientError, UnsupportedError
from .sampling import draw_sobol_samples

# SciPy >= 1.15 task codes of setulb (scipy/optimize/_lbfgsb_py.py)
_TASK_FG, _TASK_NEW_X, _TASK_CONV, _TASK_STOP = 3, 1, 4, 5


class _LbfgsbState:
    """One restart's SciPy L-BFGS-B work arrays."""

    def __init__(self, x0: np.ndarray, lb: np.ndarray, ub: np.ndarray, m: int):
        n = x0.size
        self.x = np.array(x0, dtype=np.float64)
        self.low = np.where(np.isfinite(lb), lb, 0.0).astype(np.float64)
        self.up = np.where(np.isfinite(ub), ub, 0.0).astype(np.float64)
        has_l, has_u = np.isfinite(lb), np.isfinite(ub)
        self.nbd = np.where(has_l & has_u, 2, np.where(has_l, 1, np.where(has_u, 3, 0))).astype(
            np.int32)
        self.f = np.array(0.0, dtype=np.float64)
        self.g = np.zeros(n, dtype=np.float64)
        self.wa = np.zeros(2 * m * n + 5 * n + 11 * m * m + 8 * m, dtype=np.float64)
        self.iwa = np.zeros(3 * n, dtype=np.int32)
        self.task = np.zeros(2, dtype=np.int32)
        self.ln_task = np.zeros(2, dtype=np.int32)
        self.lsave = np.zeros(4, dtype=np.int32)
        self.isave = np.zeros(44, dtype=np.int32)
        self.dsave = np.zeros(29, dtype=np.float64)
        self.nit = 0
        self.nfev = 0
        self.done = False
        self.warnflag = 0


def fmin_l_bfgs_b_batched(func, x0: np.ndarray, bounds: tuple[np.ndarray, np.ndarray],
                          maxcor: int = 10, factr: float = 1e7, pgtol: float = 1e-5,
                          maxiter: int = 15000, maxls: int = 20, maxfun: int = 15000,
                          callback=None):
    """Minimise b independent problems in lock step.  ``func(x_active, idx)``
    returns (f: (k,), g: (k, n)) for the rows ``idx`` of the batch.
    Returns (x: (b, n), f: (b,), info list)."""
    b, n = x0.shape
    lb, ub = bounds
    states = [_LbfgsbState(x0[i], lb, ub, maxcor) for i in range(b)]
    while True:
        want = []
        for i, s in enumerate(states):
            if s.done:
                continue
            while True:
                _lbfgsb.setulb(maxcor, s.x, s.low, s.up, s.nbd, s.f, s.g, factr, pgtol, s.wa,
                               s.iwa, s.task, s.lsave, s.isave, s.dsave, maxls, s.ln_task)
                t = int(s.task[0])
                if t == _TASK_FG:
                    if s.nfev >= maxfun:
                        s.task[0], s.task[1] = _TASK_STOP, 502
                        s.done, s.warnflag = True, 1
                        break
                    want.append(i)
                    break
                if t == _TASK_NEW_X:
                    s.nit += 1
                    if callback is not None:
                        callback(i, s.x)
                    if s.nit >= maxiter:
                        s.task[0], s.task[1] = _TASK_STOP, 504
                        s.done, s.warnflag = True, 1
                        break
                    continue
                s.done = True
                s.warnflag = 0 if t == _TASK_CONV else 2
                break
        if not want:
            break
        xs = np.stack([states[i].x for i in want])
        fv, gv = func(xs, want)
        for k, i in enumerate(want):
            states[i].f = np.array(float(fv[k]), dtype=np.float64)
            states[i].g = np.asarray(gv[k], dtype=np.float64).copy()
            states[i].nfev += 1
    x = np.stack([s.x for s in states])
    f = np.array([float(s.f) for s in states])
    info = [{"nit": s.nit, "nfev": s.nfev, "warnflag": s.warnflag} for s in states]
    return x, f, info


def gen_candidates_scipy(initial_conditions: Tensor, acquisition_function,
                         lower_bounds: Tensor | float | None = None,
                         upper_bounds: Tensor | float | None = None,
                         options: dict | None = None, fixed_features=None, timeout_sec=None):
    """Batched L-BFGS-B from ``initial_conditions`` (b x q x d); returns the
    candidates and their acquisition values (generation/gen.py:62-417)."""
    if fixed_features:
        raise UnsupportedError("fixed_features are not supported")
    options = options or {}
    X0 = initial_conditions
    b, q, d = X0.shape
    dev = X0.device
    nvar = q * d

    def _expand(bnd, default):
        if bnd is None:
            return np.full(nvar, default)
        t = torch.as_tensor(bnd, dtype=torch.float64).cpu()
        return np.broadcast_to(t.expand(q, d).reshape(-1).numpy() if t.dim() else
                               np.full(nvar, float(t)), (nvar,)).copy()

    lb = _expand(lower_bounds, -np.inf)
    ub = _expand(upper_bounds, np.inf)
    x0 = X0.detach().reshape(b, nvar).cpu().numpy().astype(np.float64)
    x0 = np.clip(x0, lb, ub)
    nevals = [0]

    def f_np(x: np.ndarray, idx: list[int]):
        # generation/gen.py:423-485: numpy -> device, -acqf value and gradient
        if np.isnan(x).any():
            raise RuntimeError(f"{np.isnan(x).sum()} elements of `x` are NaN.")
        X = torch.from_numpy(x).to(device=dev, dtype=torch.float64).view(-1, q, d)
        val, grad = acquisition_function.value_and_grad(X)
        nevals[0] += 1
        g = (-grad).reshape(len(idx), nvar).cpu().numpy()
        if np.isnan(g).any():
            raise OptimizationGradientError(
                f"{np.isnan(g).sum()} elements of the gradient array are NaN.", current_x=x)
        return (-val).cpu().numpy(), g

    x, f, info = fmin_l_bfgs_b_batched(
        f_np, x0, (lb, ub), maxcor=options.get("maxcor", 10),
        factr=options.get("factr", 1e7) if "ftol" not in options else
        options["ftol"] / np.finfo(float).eps,
        pgtol=options.get("gtol", options.get("pgtol", 1e-5)),
        maxiter=options.get("maxiter", 2000), maxls=options.get("maxls", 20),
        maxfun=options.get("maxfun", 15000))
    cands = torch.from_numpy(x).to(device=dev, dtype=X0.dtype).view(b, q, d)
    with torch.no_grad():
        vals = acquisition_function(cands)
    if any(i["warnflag"] == 1 for i in info):
        warnings.warn("Optimization stopped at maxiter/maxfun for some restarts.",
                      RuntimeWarning)
    return cands, vals


def initialize_q_batch(X: Tensor, acq_vals: Tensor, n: int, eta: float = 1.0,
                       generator: torch.Generator | None = None):
    """Boltzmann selection of n of the raw samples (optim/initializers.py:959)."""
    n_samples = X.shape[0]
    if n > n_samples:
        raise RuntimeError(f"n ({n}) cannot be larger than the number of provided samples "
                           f"({n_samples})")
    if n == n_samples:
        return X, acq_vals
    ystd = acq_vals.std(dim=0)
    if ystd == 0 or not torch.isfinite(ystd):
        warnings.warn("All acquisition values for raw samples points are the same or not "
                      "finite. Choosing initial conditions at random.", RuntimeWarning)
        idcs = torch.randperm(n_samples, device=X.device)[:n]
        return X[idcs], acq_vals[idcs]
    z = eta * (acq_vals - acq_vals.mean()) / ystd
    w = torch.exp(z - z.max())
    idcs = torch.multinomial(w.cpu(), n, replacement=False, generator=generator).to(X.device)
    max_idx = torch.argmax(acq_vals)
    if not (idcs == max_idx).any():
        idcs[-1] = max_idx
    return X[idcs], acq_vals[idcs]


def gen_batch_initial_conditions(acq_function, bounds: Tensor, q: int, num_restarts: int,
                                 raw_samples: int, options: dict | None = None) -> Tensor:
    """Sobol raw points -> forward-only acquisition values -> Boltzmann pick
    (optim/initializers.py:258-483)."""
    options = options or {}
    seed = options.get("seed")
    batch_limit = options.get("init_batch_limit", options.get("batch_limit", 1024))
    X_raw = draw_sobol_samples(bounds=bounds, n=raw_samples, q=q, seed=seed)
    vals = []
    with torch.no_grad():
        for start in range(0, raw_samples, batch_limit):
            vals.append(acq_function(X_raw[start:start + batch_limit]))
    Y_raw = torch.cat(vals)
    gen = torch.Generator().manual_seed(seed) if seed is not None else None
    X_init, _ = initialize_q_batch(X_raw, Y_raw, num_restarts, eta=options.get("eta", 1.0),
                                   generator=gen)
    return X_init


def get_best_candidates(batch_candidates: Tensor, batch_values: Tensor) -> Tensor:
    """generation/gen.py:680-709."""
    best = torch.argmax(batch_values.view(-1), dim=0)
    return batch_candidates[best]


def optimize_acqf(acq_function, bounds: Tensor, q: int, num_restarts: int,
                  raw_samples: int | None = None, options: dict | None = None,
                  inequality_constraints=None, equality_constraints=None,
                  nonlinear_inequality_constraints=None, fixed_features=None,
                  post_processing_func=None, batch_initial_conditions: Tensor | None = None,
                  return_best_only: bool = True, sequential: bool = False, **kwargs):
    """optim/optimize.py:619-845 for box bounds: returns (candidates, values)."""
    if inequality_constraints or equality_constraints or nonlinear_inequality_constraints:
        raise UnsupportedError("parameter constraints are not supported")
    if sequential:
        raise UnsupportedError("sequential greedy optimisation is not supported")
    options = dict(options or {})
    bounds = bounds.to(torch.float64)
    if batch_initial_conditions is None:
        if raw_samples is None:
            raise ValueError("Must specify `raw_samples` when `batch_initial_conditions` is None`.")
        batch_initial_conditions = gen_batch_initial_conditions(
            acq_function, bounds.to(acq_function.X_baseline.device), q, num_restarts,
            raw_samples, options)
    batch_limit = options.get("batch_limit", num_restarts)
    cands, vals = [], []
    for start in range(0, batch_initial_conditions.shape[0], batch_limit):
        c, v = gen_candidates_scipy(batch_initial_conditions[start:start + batch_limit],
                                    acq_function, lower_bounds=bounds[0], upper_bounds=bounds[1],
                                    options=options, fixed_features=fixed_features)
        cands.append(c)
        vals.append(v)
    batch_candidates = torch.cat(cands)
    batch_values = torch.cat(vals)
    if post_processing_func is not None:
        batch_candidates = post_processing_func(batch_candidates)
    if return_best_only:
        best = torch.argmax(batch_values.view(-1))
        return batch_candidates[best], batch_values[best]
    return batch_candidates, batch_values
