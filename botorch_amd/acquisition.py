"""qLogNoisyExpectedImprovement on libbotorch_amd.

Mirrors acquisition/logei.py:246-585 (qLogNoisyExpectedImprovement with
cache_root=True, incremental, fat=True), acquisition/monte_carlo.py:146-348
(SampleReducingMCAcquisitionFunction.forward: fat-max over q, logmeanexp over
samples), acquisition/cached_cholesky.py:63-192 (cached baseline root and the
split-seed base samples) and acquisition/utils.py:347-437
(prune_inferior_points).

``forward(X)`` runs the whole value (and, when X requires grad, the d/dX
adjoint) in one library call; the autograd Function hands the precomputed
gradient back in ``backward``.  ``value_and_grad`` is the optimiser entry.
"""
from __future__ import annotations

import ctypes
import math
import warnings

import torch
from torch import Tensor

from . import _lib
from .exceptions import BotorchWarning, NanError, NotPSDError, UnsupportedError
from .gp import SingleTaskGP
from .sampling import SobolQMCNormalSampler, draw_sobol_normal_samples

TAU_RELU = 1e-6  # acquisition/logei.py:67
TAU_MAX = 1e-2  # acquisition/logei.py:68

# "fp64": everything FP64; "mixed": the two L^-1 contractions on split-FP16 MFMA
# with FP64 segment sums; "mixed_f32": the contractions on FP32 MFMA with FP64
# segment sums.  Factors, kernel values, [alpha | A] products and the MC stage are
# FP64 in every mode.
_PRECISIONS = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED,
               "mixed_f32": _lib.PREC_MIXED_F32}


def prune_inferior_points(model: SingleTaskGP, X: Tensor, objective=None,
                          posterior_transform=None, constraints=None, num_samples: int = 2048,
                          max_frac: float = 1.0, sampler=None, marginalize_dim=None) -> Tensor:
    """Keep the points of X that are the argmax of at least one posterior sample
    (acquisition/utils.py:347-437)."""
    if objective is not None or posterior_transform is not None or constraints is not None:
        raise UnsupportedError("objectives, posterior transforms and constraints are not supported")
    if X.ndim > 2:
        raise UnsupportedError("Batched inputs `X` are currently unsupported by "
                               "`prune_inferior_points`")
    if X.size(-2) == 0:
        raise ValueError("X must have at least one point.")
    if max_frac <= 0 or max_frac > 1.0:
        raise ValueError(f"max_frac must take values in (0, 1], is {max_frac}")
    max_points = math.ceil(max_frac * X.size(-2))
    with torch.no_grad():
        posterior = model.posterior(X)
        if sampler is None:
            sampler = SobolQMCNormalSampler(sample_shape=torch.Size([num_samples]))
        samples = sampler(posterior)
    obj_vals = samples.squeeze(-1)
    is_best = torch.argmax(obj_vals, dim=-1)
    idcs, counts = torch.unique(is_best, return_counts=True)
    if len(idcs) > max_points:
        counts, order_idcs = torch.sort(counts, stable=True, descending=True)
        idcs = order_idcs[:max_points]
    return X[idcs]


class _QLogNEIFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X: Tensor, acqf: "qLogNoisyExpectedImprovement", need_grad: bool):
        val, grad = acqf._run(X.detach(), need_grad)
        ctx.save_for_backward(grad if grad is not None else torch.empty(0))
        ctx.has_grad = grad is not None
        return val

    @staticmethod
    def backward(ctx, g_out: Tensor):
        (grad,) = ctx.saved_tensors
        if not ctx.has_grad:
            raise RuntimeError("qLogNEI forward ran without gradient")
        return g_out.reshape(-1, 1, 1) * grad, None, None


class qLogNoisyExpectedImprovement(torch.nn.Module):
    """MC-based batch Log Noisy Expected Improvement (acquisition/logei.py:246)."""

    def __init__(self, model: SingleTaskGP, X_baseline: Tensor, sampler=None, objective=None,
                 posterior_transform=None, X_pending=None, constraints=None, eta=1e-3,
                 fat: bool = True, prune_baseline: bool = True, cache_root: bool | None = None,
                 tau_max: float = TAU_MAX, tau_relu: float = TAU_RELU, marginalize_dim=None,
                 incremental: bool = True, precision: str = "fp64",
                 max_workspace_bytes: int | None = None):
        super().__init__()
        if objective is not None or posterior_transform is not None or constraints is not None:
            raise UnsupportedError("objectives, posterior transforms and constraints are "
                                   "outside the accelerated path")
        if X_pending is not None:
            raise UnsupportedError("X_pending is not supported")
        if not fat:
            raise UnsupportedError("only fat=True (fat-max / log-fatplus) is implemented")
        if cache_root is False:
            raise UnsupportedError("only the cached-root path is implemented")
        if precision not in _PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}")
        for name, tau in (("tau_max", tau_max), ("tau_relu", tau_relu)):
            if not tau > 0:
                raise ValueError(f"{name} is non-positive: {tau=}.")
        self.model = model
        self.sampler = sampler or SobolQMCNormalSampler(sample_shape=torch.Size([512]))
        if len(self.sampler.sample_shape) != 1:
            raise UnsupportedError("one sample dimension is supported")
        self.tau_max = float(tau_max)
        self.tau_relu = float(tau_relu)
        self.precision = precision
        # A batch whose step workspace (the K_x, V and U rows of all its restarts)
        # exceeds this many bytes runs in restart chunks on one workspace: the rows
        # are recomputed chunk by chunk instead of being held for the whole batch.
        self.max_workspace_bytes = max_workspace_bytes
        self.prune_baseline = prune_baseline
        self.X_pending = None
        if prune_baseline:
            X_baseline = prune_inferior_points(model=model, X=X_baseline)
        self._init_baseline(X_baseline)

    # acquisition/logei.py:393-459
    def _init_baseline(self, X_baseline: Tensor) -> None:
        L = _lib.lib()
        model = self.model
        cache = model.prediction_cache()
        m = cache["model"]
        dev = cache["xs"].device
        Xb = X_baseline.to(device=dev, dtype=torch.float64).contiguous()
        n_b, d = Xb.shape
        S = int(self.sampler.sample_shape.numel())
        n_pad = cache["n_pad"]
        st = _lib.stream()
        xb_s = torch.empty(n_b, d, dtype=torch.float64, device=dev)
        _lib.check(L.bta_scale_inputs(_lib.ptr(Xb), n_b, d, _lib.ptr(cache["center"]),
                                      _lib.ptr(cache["lengthscale"]), None, _lib.ptr(xb_s), n_b,
                                      st), "scale_inputs")
        if self.precision == "mixed":  # hi/lo FP16 planes of c L^-1 and of c L^-T
            linv = torch.empty(4, n_pad, n_pad, dtype=torch.float16, device=dev)
        elif self.precision == "mixed_f32":
            linv = torch.empty(n_pad, n_pad, dtype=torch.float32, device=dev)
        else:
            linv = cache["Linv"]
        ext = torch.empty(n_pad, n_b + 1, dtype=torch.float64, device=dev)
        Lb = torch.empty(n_b, n_b, dtype=torch.float64, device=dev)
        best_f = torch.empty(S, dtype=torch.float64, device=dev)
        seed = self.sampler.seed
        Zb = draw_sobol_normal_samples(n_b, S, device=dev, seed=seed).contiguous()
        state = _lib.QLogNEIState()
        state.precision = _PRECISIONS[self.precision]
        state.q = 1
        state.n_b = n_b
        state.S = S
        state.xb_s = _lib.ptr(xb_s)
        state.linv = _lib.ptr(linv)
        state.ext = _lib.ptr(ext)
        state.Lb = _lib.ptr(Lb)
        state.best_f = _lib.ptr(best_f)
        state.tau_relu = self.tau_relu
        state.tau_max = self.tau_max
        state.max_tries = 6
        state.jitter = 1e-8
        ws = _lib.workspace(L.bta_qlognei_prepare_workspace_size(ctypes.byref(m), n_b, S), dev)
        jit = ctypes.c_double(0.0)
        lscale = ctypes.c_float(1.0)
        _lib.check(L.bta_qlognei_prepare(ctypes.byref(m), ctypes.byref(state), _lib.ptr(Zb),
                                         _lib.ptr(linv) if self.precision != "fp64" else None,
                                         _lib.ptr(ext), _lib.ptr(Lb), _lib.ptr(best_f),
                                         ctypes.byref(jit), ctypes.byref(lscale), _lib.ptr(ws),
                                         ws.numel(), st),
                   "qlognei_prepare")
        state.linv_scale = lscale.value
        self.X_baseline = Xb
        self._n_b = n_b
        self._S = S
        self._state = state
        self._keep = {"xb_s": xb_s, "linv": linv, "ext": ext, "Lb": Lb, "best_f": best_f,
                      "Zb": Zb, "cache": cache}
        self._Zt = {}
        self._ws = None
        self._baseline_L = Lb
        self._baseline_best_f = best_f
        self._levels = None  # joint-sampling fallback factors, made on first use

    def base_samples_t(self, q: int) -> Tensor:
        """(n_b + q) x S base samples: baseline columns from the baseline sampler,
        the new q columns from Sobol(n_b + q) with the same seed
        (sampling/normal.py:68-135)."""
        if q not in self._Zt:
            dev = self.X_baseline.device
            Z = draw_sobol_normal_samples(self._n_b + q, self._S, device=dev,
                                          seed=self.sampler.seed)
            Z[:, : self._n_b] = self._keep["Zb"]
            self._Zt[q] = Z.t().contiguous()
        return self._Zt[q]

    def _workspace(self, b: int, q: int) -> Tensor:
        L = _lib.lib()
        self._state.q = q
        need = L.bta_qlognei_workspace_size(ctypes.byref(self._keep["cache"]["model"]),
                                            ctypes.byref(self._state), b)
        if self._ws is None or self._ws.numel() < need:
            self._ws = _lib.workspace(need, self.X_baseline.device)
        return self._ws

    def _joint_levels(self):
        """The baseline factors of the joint-sampling fallback: level t =
        chol(Sigma_bb + j_t I), j_0 = 0, j_t = 1e-8 10^(t-1), for t = 0..max_tries,
        with Sigma_bb the model's posterior covariance at X_baseline (recomputed,
        not the cached root), and the mask of the levels whose Cholesky exists."""
        if getattr(self, "_levels", None) is None:
            from .gp import psd_safe_cholesky
            with torch.no_grad():
                cov = self.model.posterior(self.X_baseline).covariance_matrix
            if torch.isnan(cov).any():
                raise NanError("cholesky input contains NaNs")
            n_b = cov.shape[-1]
            tries = self._state.max_tries
            levels = torch.zeros(tries + 1, n_b, n_b, dtype=torch.float64, device=cov.device)
            mask = 0
            eye = torch.eye(n_b, dtype=torch.float64, device=cov.device)
            for t in range(tries + 1):
                jit = 0.0 if t == 0 else self._state.jitter * 10.0 ** (t - 1)
                try:
                    levels[t] = psd_safe_cholesky(cov + jit * eye, max_tries=0)
                    mask |= 1 << t
                except NotPSDError:
                    pass
            self._levels = (levels, mask)
        return self._levels

    def _chunk(self, b: int, q: int) -> int:
        """Restarts per library call: all b, or the most whose workspace fits in
        max_workspace_bytes."""
        if self.max_workspace_bytes is None:
            return b
        L = _lib.lib()
        self._state.q = q
        m = ctypes.byref(self._keep["cache"]["model"])

        def need(k):
            return L.bta_qlognei_workspace_size(m, ctypes.byref(self._state), k)

        if need(b) <= self.max_workspace_bytes:
            return b
        lo, hi = 1, b  # need(lo) fits (or nothing does), need(hi) does not
        if need(1) > self.max_workspace_bytes:
            raise UnsupportedError(f"max_workspace_bytes={self.max_workspace_bytes} is below "
                                   f"one restart's workspace ({need(1)} bytes)")
        while hi - lo > 1:
            mid = (lo + hi) // 2
            if need(mid) <= self.max_workspace_bytes:
                lo = mid
            else:
                hi = mid
        return lo

    def _launch(self, X: Tensor, need_grad: bool, joint: bool = False):
        """Enqueue one value(+grad) evaluation; returns device tensors
        (value b, grad b x q x d or None, status b).  joint: sample from the
        Cholesky of the joint (X_baseline, X) posterior instead of the cached
        baseline root."""
        L = _lib.lib()
        b, q, d = X.shape
        if q > 8:
            raise UnsupportedError("q <= 8 is supported")
        Xc = X.to(device=self.X_baseline.device, dtype=torch.float64).contiguous()
        Zt = self.base_samples_t(q)
        chunk = self._chunk(b, q)
        ws = self._workspace(chunk, q)
        st = self._state
        st.q = q
        st.Zt = _lib.ptr(Zt)
        value = torch.empty(b, dtype=torch.float64, device=Xc.device)
        grad = torch.empty(b, q, d, dtype=torch.float64, device=Xc.device) if need_grad else None
        status = torch.empty(b, dtype=torch.int32, device=Xc.device)
        if joint:
            levels, mask = self._joint_levels()
            st.Lb_levels = _lib.ptr(levels)
            st.lb_levels_mask = mask
        try:
            for r0 in range(0, b, chunk):
                r1 = min(b, r0 + chunk)
                _lib.check(L.bta_qlognei_value_grad(
                    ctypes.byref(self._keep["cache"]["model"]), ctypes.byref(st),
                    _lib.ptr(Xc[r0:r1]), r1 - r0, _lib.ptr(value[r0:r1]),
                    None if grad is None else _lib.ptr(grad[r0:r1]), _lib.ptr(status[r0:r1]),
                    _lib.ptr(ws), ws.numel(), _lib.stream()), "qlognei_value_grad")
        finally:
            st.Lb_levels = None
            st.lb_levels_mask = 0
        return value, grad, status

    @staticmethod
    def check_status(status: Tensor) -> None:
        """Raise like the reference would on failed samples (utils/low_rank.py:162-171,
        psd_safe_cholesky NotPSDError); warn when jitter was added."""
        s = torch.zeros(3, dtype=torch.bool, device=status.device)
        for i, f in enumerate((_lib.STATUS_NAN, _lib.STATUS_NOT_PSD, _lib.STATUS_JITTER)):
            s[i] = ((status & f) != 0).any()
        nan, notpsd, jit = s.tolist()
        if nan:
            raise NanError("Samples contain nans or infs.")
        if notpsd:
            raise NotPSDError("Matrix not positive definite after repeatedly adding jitter.")
        if jit:
            warnings.warn("A not p.d., added jitter to the diagonal", RuntimeWarning)

    def autotune(self, X: Tensor) -> dict:
        """Times the plan candidates of the step's two split-FP16 products on X
        (b x q x d) and keeps the fastest for this shape (bta_qlognei_autotune);
        returns the chosen (tile width, workers, XCD groups) of both.  Results do
        not change: every plan computes each output element in the same order."""
        L = _lib.lib()
        b, q, d = X.shape
        Xc = X.to(device=self.X_baseline.device, dtype=torch.float64).contiguous()
        st = self._state
        st.q = q
        st.Zt = _lib.ptr(self.base_samples_t(q))
        ws = self._workspace(b, q)
        value = torch.empty(b, dtype=torch.float64, device=Xc.device)
        grad = torch.empty(b, q, d, dtype=torch.float64, device=Xc.device)
        status = torch.empty(b, dtype=torch.int32, device=Xc.device)
        chosen = (ctypes.c_int32 * 6)()
        _lib.check(L.bta_qlognei_autotune(ctypes.byref(self._keep["cache"]["model"]),
                                          ctypes.byref(st), _lib.ptr(Xc), b, _lib.ptr(value),
                                          _lib.ptr(grad), _lib.ptr(status), _lib.ptr(ws),
                                          ws.numel(), chosen, _lib.stream()), "qlognei_autotune")
        c = list(chosen)
        return {"forward": {"bn": c[0], "workers": c[1], "xcd_groups": c[2]},
                "adjoint": {"bn": c[3], "workers": c[4], "xcd_groups": c[5]}}

    @staticmethod
    def _cached_root_failed(status: Tensor) -> bool:
        return bool(((status & (_lib.STATUS_NAN | _lib.STATUS_NOT_PSD)) != 0).any())

    def _run(self, X: Tensor, need_grad: bool):
        value, grad, status = self._launch(X, need_grad)
        if self._cached_root_failed(status):
            # acquisition/cached_cholesky.py:143-158: the low-rank update raised
            # NanError / NotPSDError, so the whole batch is sampled again from the
            # joint posterior's Cholesky
            warnings.warn("Low-rank cholesky updates failed due NaNs or due to an "
                          "ill-conditioned covariance matrix. Falling back to standard "
                          "sampling.", BotorchWarning, stacklevel=3)
            value, grad, status = self._launch(X, need_grad, joint=True)
        self.check_status(status)
        return value, grad

    def value_and_grad(self, X: Tensor):
        """Batched value and d value[b] / d X[b] for X: b x q x d (one L-BFGS step)."""
        Xt = X if X.dim() > 2 else X.unsqueeze(0)
        value, grad = self._run(Xt.detach(), True)
        return value, grad

    def forward(self, X: Tensor) -> Tensor:
        """SampleReducingMCAcquisitionFunction.forward with t_batch_mode_transform
        (utils/transforms.py:239-323): X is batch_shape x q x d."""
        if X.dim() < 2:
            raise ValueError(f"{type(self).__name__} requires X to have at least 2 dimensions,"
                             f" but received X with only {X.dim()} dimensions.")
        Xt = X if X.dim() > 2 else X.unsqueeze(0)
        batch_shape = Xt.shape[:-2]
        q, d = Xt.shape[-2:]
        Xf = Xt.reshape(-1, q, d)
        need_grad = torch.is_grad_enabled() and Xf.requires_grad
        if need_grad:
            val = _QLogNEIFn.apply(Xf, self, True)
        else:
            val, _ = self._run(Xf.detach(), False)
        return val.reshape(batch_shape).to(X.dtype)


class BatchedQLogNEI:
    """qLogNEI value+grad of many independent models in one batched step.

    Stands in for evaluating the acquisition functions of many SingleTaskGPs
    one after another (one L-BFGS-B step of each model's optimize_acqf,
    generation/gen.py:423-485): the step over all of them is five launches, one
    per stage, each covering every model (``bta_qlognei_batch_value_grad``).
    Each acquisition function keeps its own model, baseline and base samples;
    the batch shares q, the kernel kind and the mixed precision.  The query
    points live in fixed device buffers (``X[i]``, restarts[i] x q x d) that the
    step reads in place, so a captured HIP graph of ``launch`` stays valid when
    they are refilled.
    """

    def __init__(self, acqfs, restarts, q: int, need_grad: bool = True):
        L = _lib.lib()
        if isinstance(restarts, int):
            restarts = [restarts] * len(acqfs)
        if len(acqfs) != len(restarts) or not acqfs:
            raise ValueError("one restart count per acquisition function")
        self.acqfs = list(acqfs)
        self.q = q
        self.need_grad = need_grad
        dev = self.acqfs[0].X_baseline.device
        self.X, self.value, self.grad, self.status = [], [], [], []
        self._states = []
        items = (_lib.QLogNEIBatchItem * len(acqfs))()
        for i, (acq, b) in enumerate(zip(self.acqfs, restarts)):
            if acq.precision != "mixed":
                raise UnsupportedError("the batched step runs the mixed precision only")
            d = acq.X_baseline.shape[-1]
            self.X.append(torch.zeros(b, q, d, dtype=torch.float64, device=dev))
            self.value.append(torch.empty(b, dtype=torch.float64, device=dev))
            self.grad.append(torch.empty(b, q, d, dtype=torch.float64, device=dev)
                             if need_grad else None)
            self.status.append(torch.empty(b, dtype=torch.int32, device=dev))
            st = _lib.QLogNEIState()  # a copy: this batch's q and base samples
            ctypes.memmove(ctypes.byref(st), ctypes.byref(acq._state), ctypes.sizeof(st))
            st.q = q
            st.Zt = _lib.ptr(acq.base_samples_t(q))
            self._states.append(st)
            it = items[i]
            it.model = ctypes.pointer(acq._keep["cache"]["model"])
            it.state = ctypes.pointer(st)
            it.X = _lib.ptr(self.X[i])
            it.b = b
            it.value = _lib.ptr(self.value[i])
            it.grad = _lib.ptr(self.grad[i])
            it.status = _lib.ptr(self.status[i])
        self._items = items
        need = L.bta_qlognei_batch_workspace_size(items, len(acqfs))
        if need == 0:
            raise ValueError("invalid batch items")
        self.ws = _lib.workspace(need, dev)
        self.plan = _lib.QLogNEIBatchPlan()
        _lib.check(L.bta_qlognei_batch_prepare(items, len(acqfs), _lib.ptr(self.ws),
                                               self.ws.numel(), ctypes.byref(self.plan),
                                               _lib.stream()), "qlognei_batch_prepare")

    def launch(self) -> None:
        """Enqueue one step over every model (reads ``X``, writes ``value``,
        ``grad``, ``status``); no host synchronisation."""
        _lib.check(_lib.lib().bta_qlognei_batch_value_grad(ctypes.byref(self.plan),
                                                           _lib.stream()),
                   "qlognei_batch_value_grad")

    def value_and_grad(self, Xs):
        """Copies Xs[i] (restarts[i] x q x d) into the buffers, runs the step and
        returns ([value_i], [grad_i]), raising like the one-model step on failed
        samples."""
        for buf, X in zip(self.X, Xs):
            buf.copy_(X)
        self.launch()
        for i, acq in enumerate(self.acqfs):
            if qLogNoisyExpectedImprovement._cached_root_failed(self.status[i]):
                # this model's cached-root update failed: its acquisition function
                # samples it again from the joint posterior (and warns)
                v, g = acq._run(self.X[i], self.need_grad)
                self.value[i].copy_(v)
                if self.need_grad:
                    self.grad[i].copy_(g)
                self.status[i].zero_()
        qLogNoisyExpectedImprovement.check_status(torch.cat(self.status))
        return ([v.clone() for v in self.value],
                [g.clone() for g in self.grad] if self.need_grad else None)
