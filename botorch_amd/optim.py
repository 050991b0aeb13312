"""optimize_acqf: multi-start L-BFGS-B over the batched qLogNEI value+grad.

Mirrors optim/optimize.py:619-845 (optimize_acqf -> _optimize_acqf_batch),
optim/initializers.py:258-483,959 (gen_batch_initial_conditions,
initialize_q_batch), generation/gen.py:62-507,680-709 (gen_candidates_scipy,
f_np_wrapper, get_best_candidates) and optim/batched_lbfgs_b.py:169-636 (lock-step
batched L-BFGS-B: one SciPy L-BFGS-B state per restart, one batched
function+gradient evaluation per step for all restarts that asked for one).

Two drivers of the same lock-step loop:
  - ``fmin_l_bfgs_b_device`` keeps every restart's L-BFGS-B state on the GPU
    (bta_lbfgsb_init / _step / _results): one round is the acquisition
    value+grad for all restarts followed by one L-BFGS-B kernel, with no host
    round trip; the host only checks every few rounds whether all restarts
    ended, and the rounds can be replayed from a HIP graph.  It is the path of
    gen_candidates_scipy for this library's acquisition functions on the GPU.
  - ``fmin_l_bfgs_b_batched`` runs SciPy's own L-BFGS-B per restart on the host
    with one batched evaluation per step, as the reference does (used for
    other callables, host callbacks, or ``options["device_lbfgsb"] = False``).
Both take the same steps per restart (the device kernel is checked against
SciPy's setulb evaluation by evaluation).
"""
from __future__ import annotations

import ctypes
import logging
import math
import time
import warnings

import numpy as np
import torch
from scipy.optimize import _lbfgsb
from torch import Tensor

from . import _lib
from .exceptions import (BotorchWarning, OptimizationGradientError, OptimizationWarning,
                         UnsupportedError)
from .sampling import draw_sobol_samples

logger = logging.getLogger(__name__)

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
                        s.message = _DEV_TASKS[5]
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
                        s.message = _DEV_TASKS[4]
                        break
                    continue
                s.done = True
                s.warnflag = 0 if t == _TASK_CONV else 2
                s.message = "CONVERGENCE" if t == _TASK_CONV else (
                    f"ABNORMAL_TERMINATION_IN_LNSRCH (setulb task {t}, {int(s.task[1])})")
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
    info = [{"nit": s.nit, "nfev": s.nfev, "warnflag": s.warnflag,
             "message": getattr(s, "message", "")} for s in states]
    return x, f, info


# task codes of bta_lbfgsb_results
_DEV_TASKS = {0: "RUNNING", 1: "CONVERGENCE: NORM_OF_PROJECTED_GRADIENT_<=_PGTOL",
              2: "CONVERGENCE: REL_REDUCTION_OF_F_<=_FACTR*EPSMCH",
              3: "ABNORMAL_TERMINATION_IN_LNSRCH", 4: "STOP: TOTAL NO. of ITERATIONS REACHED LIMIT",
              5: "STOP: TOTAL NO. of f AND g EVALUATIONS EXCEEDS LIMIT",
              6: "NaN in the gradient"}


def _nbd(lb: np.ndarray, ub: np.ndarray) -> np.ndarray:
    has_l, has_u = np.isfinite(lb), np.isfinite(ub)
    return np.where(has_l & has_u, 2, np.where(has_l, 1, np.where(has_u, 3, 0))).astype(np.int32)


def fmin_l_bfgs_b_device(evaluate, x0: Tensor, bounds: tuple[np.ndarray, np.ndarray],
                         maxcor: int = 10, factr: float = 1e7, pgtol: float = 1e-5,
                         maxiter: int = 15000, maxls: int = 20, maxfun: int = 15000,
                         fscale: float = 1.0, rounds_per_check: int = 8,
                         use_graph: bool = False, trace: list | None = None):
    """Minimise b problems in lock step with their L-BFGS-B states on the GPU.

    ``evaluate(x_eval)`` enqueues, without synchronising, the objective at the
    b x n float64 device tensor ``x_eval`` and returns (f (b), g (b x n)[,
    status (b) int32]); the minimised objective is ``fscale * f``.  Each round
    evaluates every problem once and advances each to its next point; the host
    reads the count of ended problems every ``rounds_per_check`` rounds (the
    only synchronisation), and with ``use_graph`` those rounds replay from one
    captured HIP graph.  ``trace``, when a list, receives x_eval after every
    round (one synchronisation per round; for tests).

    Returns (x: b x n, f: b (in the minimised scale), info: list of dicts with
    nit, nfev, warnflag, task, status, message)."""
    L = _lib.lib()
    b, n = x0.shape
    dev = x0.device
    lb, ub = (np.broadcast_to(np.asarray(v, dtype=np.float64), (n,)) for v in bounds)
    if (lb > ub).any():
        raise ValueError("LBFGSB - one of the lower bounds is greater than an upper bound.")
    size = L.bta_lbfgsb_state_size(n, maxcor)
    if size == 0:
        raise UnsupportedError(f"{n} variables with maxcor={maxcor} exceed the device "
                               "L-BFGS-B's per-restart LDS budget")
    nbd = torch.from_numpy(_nbd(lb, ub)).to(dev)
    lo = torch.from_numpy(np.where(np.isfinite(lb), lb, 0.0)).to(dev)
    up = torch.from_numpy(np.where(np.isfinite(ub), ub, 0.0)).to(dev)
    state = torch.empty(b * size, dtype=torch.uint8, device=dev)
    x_eval = torch.empty(b, n, dtype=torch.float64, device=dev)
    ndone = torch.zeros(1, dtype=torch.int32, device=dev)
    xs = x0.detach().to(device=dev, dtype=torch.float64).contiguous()
    _lib.check(L.bta_lbfgsb_init(_lib.ptr(state), b, n, maxcor, _lib.ptr(xs), _lib.ptr(lo),
                                 _lib.ptr(up), _lib.ptr(nbd), float(factr), float(pgtol), maxls,
                                 maxiter, maxfun, _lib.ptr(x_eval), _lib.ptr(ndone),
                                 _lib.stream()), "lbfgsb_init")

    def one_round():
        out = evaluate(x_eval)
        f, g = out[0], out[1]
        st = out[2] if len(out) > 2 else None
        g = g.reshape(b, n)
        if f.dtype != torch.float64 or g.dtype != torch.float64:
            raise TypeError("evaluate must return float64 f and g")
        _lib.check(L.bta_lbfgsb_step(_lib.ptr(state), b, n, maxcor, _lib.ptr(lo), _lib.ptr(up),
                                     _lib.ptr(nbd), _lib.ptr(f.contiguous()),
                                     _lib.ptr(g.contiguous()), float(fscale),
                                     None if st is None else _lib.ptr(st.contiguous()),
                                     _lib.ptr(x_eval), _lib.ptr(ndone), _lib.stream()),
                   "lbfgsb_step")

    # every round ends a problem or spends one of its evaluations, so all end
    # within maxfun + 1 rounds
    max_rounds = maxfun + 2
    rounds = 0
    if trace is not None:
        trace.append(x_eval.clone())
        while int(ndone.item()) < b:
            one_round()
            rounds += 1
            trace.append(x_eval.clone())
            if rounds > max_rounds:
                raise RuntimeError("device L-BFGS-B did not end within maxfun rounds")
    else:
        if use_graph and int(ndone.item()) < b:
            # one eager round first: the evaluation's workspaces and plans exist
            # before capture
            one_round()
            rounds += 1
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                for _ in range(rounds_per_check):
                    one_round()
            run = graph.replay
        else:
            def run():
                for _ in range(rounds_per_check):
                    one_round()
        while int(ndone.item()) < b:
            run()
            rounds += rounds_per_check
            if rounds > max_rounds + rounds_per_check:
                raise RuntimeError("device L-BFGS-B did not end within maxfun rounds")
    x = torch.empty(b, n, dtype=torch.float64, device=dev)
    f = torch.empty(b, dtype=torch.float64, device=dev)
    info = torch.empty(b, 6, dtype=torch.int32, device=dev)
    _lib.check(L.bta_lbfgsb_results(_lib.ptr(state), b, n, maxcor, _lib.ptr(x), _lib.ptr(f),
                                    _lib.ptr(info), _lib.stream()), "lbfgsb_results")
    rows = info.cpu().tolist()
    infos = [{"nit": r[0], "nfev": r[1], "warnflag": r[2], "task": r[3], "status": r[4],
              "updates": r[5], "message": _DEV_TASKS.get(r[3], str(r[3]))} for r in rows]
    return x, f, infos


def _process_result(info: dict, options: dict) -> None:
    """generation/gen.py:711-746: iteration and evaluation limits are logged,
    other failures warn with OptimizationWarning."""
    if info["warnflag"] == 0:
        return
    msg = info.get("message", "")
    if "ITERATIONS REACHED LIMIT" in msg:
        logger.info("`scipy.optimize.minimize` exited by reaching the iteration limit of "
                    f"`maxiter: {options.get('maxiter')}`.")
    elif "EVALUATIONS EXCEEDS LIMIT" in msg:
        logger.info("`scipy.optimize.minimize` exited by reaching the function evaluation "
                    f"limit of `maxfun: {options.get('maxfun')}`.")
    else:
        with warnings.catch_warnings():
            warnings.simplefilter("always", category=OptimizationWarning)
            warnings.warn(f"Optimization failed within `scipy.optimize.minimize` with status "
                          f"{info['warnflag']} and message {msg}.", OptimizationWarning,
                          stacklevel=3)


def _device_path(acquisition_function, X0: Tensor, options: dict) -> bool:
    return (options.get("device_lbfgsb", True) and X0.is_cuda
            and hasattr(acquisition_function, "_launch") and options.get("callback") is None)


def gen_candidates_scipy(initial_conditions: Tensor, acquisition_function,
                         lower_bounds: Tensor | float | None = None,
                         upper_bounds: Tensor | float | None = None,
                         options: dict | None = None, fixed_features=None, timeout_sec=None):
    """Batched L-BFGS-B from ``initial_conditions`` (b x q x d); returns the
    candidates and their acquisition values (generation/gen.py:62-417).

    For this library's acquisition functions on the GPU the optimiser runs on
    the device (fmin_l_bfgs_b_device; ``options["device_lbfgsb"] = False``
    selects the host SciPy loop, ``options["lbfgsb_graph"]`` replays the rounds
    from a HIP graph).  Results of each restart are processed as
    generation/gen.py:711-746 does: limits are logged, other failures raise an
    OptimizationWarning."""
    if fixed_features:
        raise UnsupportedError("fixed_features are not supported")
    options = options or {}
    X0 = initial_conditions
    if X0.dim() == 2:
        X0 = X0.unsqueeze(0)
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
    lbfgs_kw = dict(maxcor=options.get("maxcor", 10),
                    factr=options.get("factr", 1e7) if "ftol" not in options else
                    options["ftol"] / np.finfo(float).eps,
                    pgtol=options.get("gtol", options.get("pgtol", 1e-5)),
                    maxiter=options.get("maxiter", 2000), maxls=options.get("maxls", 20),
                    maxfun=options.get("maxfun", 15000))
    if _device_path(acquisition_function, X0, options):
        x0 = torch.maximum(torch.minimum(
            X0.detach().reshape(b, nvar).to(torch.float64),
            torch.from_numpy(ub).to(dev)), torch.from_numpy(lb).to(dev))

        def run(joint: bool):
            def evaluate(x_eval):
                value, grad, status = acquisition_function._launch(x_eval.view(b, q, d), True,
                                                                   joint=joint)
                return value, grad, status
            return fmin_l_bfgs_b_device(evaluate, x0, (lb, ub), fscale=-1.0,
                                        use_graph=options.get("lbfgsb_graph", False),
                                        **lbfgs_kw)

        x, f, info = run(False)
        status = 0
        for i in info:
            status |= i["status"]
        if status & (_lib.STATUS_NAN | _lib.STATUS_NOT_PSD):
            # acquisition/cached_cholesky.py:143-158: sample from the joint
            # posterior's Cholesky instead of the cached baseline root
            warnings.warn("Low-rank cholesky updates failed due NaNs or due to an "
                          "ill-conditioned covariance matrix. Falling back to standard "
                          "sampling.", BotorchWarning, stacklevel=2)
            x, f, info = run(True)
            status = 0
            for i in info:
                status |= i["status"]
        st = torch.tensor([status], dtype=torch.int32, device=dev)
        acquisition_function.check_status(st)
        nan_rows = [k for k, i in enumerate(info) if i["task"] == 6]
        if nan_rows:
            raise OptimizationGradientError(
                f"NaN elements in the gradient array of restarts {nan_rows[:8]}.",
                current_x=x[nan_rows].cpu().numpy())
        cands = x.to(dtype=X0.dtype).view(b, q, d)
    else:
        x0 = X0.detach().reshape(b, nvar).cpu().numpy().astype(np.float64)
        x0 = np.clip(x0, lb, ub)

        def f_np(x: np.ndarray, idx: list[int]):
            # generation/gen.py:423-485: numpy -> device, -acqf value and gradient
            if np.isnan(x).any():
                raise RuntimeError(f"{np.isnan(x).sum()} elements of `x` are NaN.")
            X = torch.from_numpy(x).to(device=dev, dtype=torch.float64).view(-1, q, d)
            val, grad = acquisition_function.value_and_grad(X)
            g = (-grad).reshape(len(idx), nvar).cpu().numpy()
            if np.isnan(g).any():
                raise OptimizationGradientError(
                    f"{np.isnan(g).sum()} elements of the gradient array are NaN.",
                    current_x=x)
            return (-val).cpu().numpy(), g

        x, f, info = fmin_l_bfgs_b_batched(f_np, x0, (lb, ub),
                                           callback=options.get("callback"), **lbfgs_kw)
        cands = torch.from_numpy(x).to(device=dev, dtype=X0.dtype).view(b, q, d)
    for i in info:
        _process_result(i, lbfgs_kw)
    # generation/gen.py:402-413: candidates clamped to the bounds (a step that
    # ends at its bound can overshoot by an ulp) and in the input's shape
    lo_t = torch.from_numpy(lb).to(device=dev, dtype=cands.dtype).view(q, d)
    hi_t = torch.from_numpy(ub).to(device=dev, dtype=cands.dtype).view(q, d)
    cands = torch.maximum(torch.minimum(cands, hi_t), lo_t).reshape(initial_conditions.shape)
    with torch.no_grad():
        vals = acquisition_function(cands)
    return cands, vals


# options consumed by the initial-condition generator (optim/optimize.py:45-60)
_INIT_OPTION_KEYS = {"alpha", "batch_limit", "eta", "init_batch_limit", "nonnegative",
                     "n_burnin", "sample_around_best", "sample_around_best_sigma",
                     "sample_around_best_prob_perturb", "seed", "thinning"}


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
                  return_best_only: bool = True, sequential: bool = False,
                  retry_on_optimization_warning: bool = True, **kwargs):
    """optim/optimize.py:619-845 for box bounds: returns (candidates, values).

    As _optimize_acqf_batch (optimize.py:364-526): provided initial conditions
    are completed by generated ones up to num_restarts, the restarts run in
    batches of batch_limit, and when a batch raised an OptimizationWarning the
    generated initial conditions are drawn again and the optimisation retried
    once."""
    if inequality_constraints or equality_constraints or nonlinear_inequality_constraints:
        raise UnsupportedError("parameter constraints are not supported")
    if sequential:
        raise UnsupportedError("sequential greedy optimisation is not supported")
    options = dict(options or {})
    bounds = bounds.to(torch.float64)
    dev = acq_function.X_baseline.device
    provided = batch_initial_conditions
    required = num_restarts
    if provided is not None and provided.dim() == 3:
        required -= provided.shape[0]

    def make_ics():
        generated = None
        if raw_samples is not None and required > 0:
            generated = gen_batch_initial_conditions(acq_function, bounds.to(dev), q, required,
                                                     raw_samples, options)
        if provided is not None and generated is not None:
            return torch.cat([provided, generated])
        if provided is not None:
            return provided
        if generated is not None:
            return generated
        raise ValueError("Either `batch_initial_conditions` or `raw_samples` must be set.")

    batch_limit = options.get("batch_limit", num_restarts)
    gen_options = {k: v for k, v in options.items() if k not in _INIT_OPTION_KEYS}

    def optimize_batches(ics):
        cands, vals, ws_all = [], [], []
        for start in range(0, ics.shape[0], batch_limit):
            with warnings.catch_warnings(record=True) as ws:
                warnings.simplefilter("always", category=OptimizationWarning)
                c, v = gen_candidates_scipy(ics[start:start + batch_limit], acq_function,
                                            lower_bounds=bounds[0], upper_bounds=bounds[1],
                                            options=gen_options, fixed_features=fixed_features)
            ws_all += ws
            cands.append(c)
            vals.append(v)
        return torch.cat(cands), torch.cat(vals).flatten(), ws_all

    # warnings raised inside gen_candidates_scipy are recorded, not shown, as in
    # optimize.py:446-463
    batch_candidates, batch_values, ws = optimize_batches(make_ics())
    if any(issubclass(w.category, OptimizationWarning) for w in ws) and \
            retry_on_optimization_warning:
        msgs = [str(w.message) for w in ws]
        if provided is not None and required <= 0:
            warnings.warn("Optimization failed in `gen_candidates_scipy` with the following "
                          f"warning(s):\n{msgs}\nBecause you specified "
                          "`batch_initial_conditions` larger than required `num_restarts`, "
                          "optimization will not be retried with new initial conditions and "
                          "will proceed with the current solution. Suggested remediation: "
                          "Try again with different `batch_initial_conditions`, don't provide "
                          "`batch_initial_conditions`, or increase `num_restarts`.",
                          RuntimeWarning, stacklevel=2)
        else:
            warnings.warn("Optimization failed in `gen_candidates_scipy` with the following "
                          f"warning(s):\n{msgs}\nTrying again with a new set of initial "
                          "conditions.", RuntimeWarning, stacklevel=2)
            if raw_samples is not None and required > 0:
                batch_candidates, batch_values, ws = optimize_batches(make_ics())
                if any(issubclass(w.category, OptimizationWarning) for w in ws):
                    warnings.warn("Optimization failed on the second try, after generating a "
                                  "new set of initial conditions.", RuntimeWarning,
                                  stacklevel=2)
    if post_processing_func is not None:
        batch_candidates = post_processing_func(batch_candidates)
        with torch.no_grad():
            batch_values = torch.cat([acq_function(c) for c in
                                      batch_candidates.split(batch_limit, dim=0)]).flatten()
    if return_best_only:
        best = torch.argmax(batch_values.view(-1))
        return batch_candidates[best], batch_values[best]
    return batch_candidates, batch_values
