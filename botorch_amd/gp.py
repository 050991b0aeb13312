"""SingleTaskGP and the modules it is assembled from, with BoTorch's interface.

Host-side mirror of botorch/models/gp_regression.py:53-214 (SingleTaskGP),
models/gpytorch.py:544-610 (posterior), models/utils/gpytorch_modules.py:74-133
(default likelihood / kernel priors and constraints) and
models/transforms/outcome.py:236-511 (Standardize), plus the gpytorch modules
those build on (MaternKernel, RBFKernel, ScaleKernel, GaussianLikelihood,
ConstantMean, LogNormalPrior, GreaterThan / Positive constraints).

The modules only hold hyperparameters.  Every kernel matrix, factorisation,
solve and posterior contraction runs in libbotorch_amd (HIP) through
``botorch_amd._lib``; there is no PyTorch fallback for that arithmetic.
"""
from __future__ import annotations

import ctypes
import math

import torch
from torch import Tensor

from . import _lib
from .exceptions import InputDataError, NanError, NotPSDError, UnsupportedError

SQRT2 = math.sqrt(2)
SQRT3 = math.sqrt(3)
MIN_INFERRED_NOISE_LEVEL = 1e-4  # models/utils/gpytorch_modules.py


# ----------------------------------------------------------------------------
# constraints and priors (gpytorch.constraints / gpytorch.priors)
# ----------------------------------------------------------------------------
class GreaterThan:
    """value = lower + softplus(raw) (transform='softplus') or value = raw
    (transform=None, bounds enforced by the optimiser)."""

    def __init__(self, lower_bound: float, transform: str | None = "softplus",
                 initial_value: float | None = None):
        self.lower_bound = float(lower_bound)
        self.transform = transform
        self.initial_value = initial_value

    def transform_value(self, raw: Tensor) -> Tensor:
        if self.transform is None:
            return raw
        return self.lower_bound + torch.nn.functional.softplus(raw)

    def inverse_transform(self, value: Tensor) -> Tensor:
        if self.transform is None:
            return value
        v = value - self.lower_bound
        return v + torch.log(-torch.expm1(-v))  # softplus^-1

    def dvalue_draw(self, raw: Tensor) -> Tensor:
        if self.transform is None:
            return torch.ones_like(raw)
        return torch.sigmoid(raw)

    def raw_bounds(self) -> tuple[float | None, float | None]:
        return (self.lower_bound, None) if self.transform is None else (None, None)


class Positive(GreaterThan):
    def __init__(self, initial_value: float | None = None):
        super().__init__(0.0, "softplus", initial_value)


class LogNormalPrior:
    """LogNormal(loc, scale) prior (torch.distributions.LogNormal)."""

    def __init__(self, loc: float, scale: float):
        self.loc = float(loc)
        self.scale = float(scale)

    @property
    def mode(self) -> float:
        return math.exp(self.loc - self.scale**2)

    def log_prob(self, x: Tensor) -> Tensor:
        lx = torch.log(x)
        return (-((lx - self.loc) ** 2) / (2 * self.scale**2) - math.log(self.scale)
                - 0.5 * math.log(2 * math.pi) - lx)

    def dlog_prob(self, x: Tensor) -> Tensor:
        lx = torch.log(x)
        return (-(lx - self.loc) / (self.scale**2) - 1.0) / x

    def sample(self, shape, generator=None) -> Tensor:
        z = torch.randn(shape, generator=generator, dtype=torch.float64)
        return torch.exp(self.loc + self.scale * z)


def _raw_zero_value(c: GreaterThan) -> float:
    """The value gpytorch starts a softplus-constrained parameter at: its raw
    parameter is initialised to zero, so the value is lower_bound + softplus(0)
    = lower_bound + ln 2."""
    return float(c.transform_value(torch.zeros((), dtype=torch.float64)))


class _HyperModule(torch.nn.Module):
    """A module with constrained scalar / vector hyperparameters."""

    def _register(self, name: str, value: Tensor, constraint: GreaterThan, prior=None):
        raw = constraint.inverse_transform(value.detach().to(torch.float64))
        self.register_parameter(f"raw_{name}", torch.nn.Parameter(raw))
        setattr(self, f"raw_{name}_constraint", constraint)
        setattr(self, f"{name}_prior", prior)

    def _value(self, name: str) -> Tensor:
        c = getattr(self, f"raw_{name}_constraint")
        return c.transform_value(getattr(self, f"raw_{name}"))

    def _set(self, name: str, value) -> None:
        raw = getattr(self, f"raw_{name}")
        c = getattr(self, f"raw_{name}_constraint")
        v = torch.as_tensor(value, dtype=raw.dtype, device=raw.device).expand_as(raw)
        if isinstance(c, GreaterThan) and c.transform is not None and bool(
                (v < c.lower_bound).any()):
            # gpytorch's Module.initialize refuses values outside a constraint
            raise RuntimeError(
                f"Attempting to manually set a parameter value that is out of bounds of its "
                f"current constraints, {type(c).__name__}({c.lower_bound:.3E}). Most likely, "
                f"you want to do the following:\n likelihood = GaussianLikelihood"
                f"(noise_constraint=gpytorch.constraints.GreaterThan(better_lower_bound))")
        with torch.no_grad():
            raw.copy_(c.inverse_transform(v))


# ----------------------------------------------------------------------------
# kernels
# ----------------------------------------------------------------------------
class _ARDKernel(_HyperModule):
    kind: int = -1

    def __init__(self, ard_num_dims: int, lengthscale_prior=None,
                 lengthscale_constraint: GreaterThan | None = None, **kwargs):
        super().__init__()
        if kwargs.get("batch_shape") not in (None, torch.Size()):
            raise UnsupportedError("batched kernels are not supported")
        if kwargs.get("active_dims") is not None:
            raise UnsupportedError("active_dims is not supported")
        self.ard_num_dims = int(ard_num_dims)
        c = lengthscale_constraint or Positive()
        init = c.initial_value
        if init is None:
            init = 1.0 if c.transform is None else _raw_zero_value(c)
            init = max(init, c.lower_bound)
        self._register("lengthscale", torch.full((self.ard_num_dims,), float(init),
                                                  dtype=torch.float64), c, lengthscale_prior)

    @property
    def lengthscale(self) -> Tensor:
        return self._value("lengthscale")

    @lengthscale.setter
    def lengthscale(self, v) -> None:
        self._set("lengthscale", v)

    @property
    def base_kernel(self):
        return self

    @property
    def outputscale(self) -> Tensor:
        return torch.ones((), dtype=torch.float64, device=self.raw_lengthscale.device)


class MaternKernel(_ARDKernel):
    """gpytorch MaternKernel; only nu=2.5 is on the accelerated path."""

    kind = _lib.KERNEL_MATERN52

    def __init__(self, nu: float = 2.5, ard_num_dims: int | None = None, **kwargs):
        if nu != 2.5:
            raise UnsupportedError("only Matern nu=2.5 is implemented")
        if ard_num_dims is None:
            raise UnsupportedError("ard_num_dims is required")
        super().__init__(ard_num_dims, **kwargs)
        self.nu = nu


class RBFKernel(_ARDKernel):
    kind = _lib.KERNEL_RBF

    def __init__(self, ard_num_dims: int | None = None, **kwargs):
        if ard_num_dims is None:
            raise UnsupportedError("ard_num_dims is required")
        super().__init__(ard_num_dims, **kwargs)


class ScaleKernel(_HyperModule):
    """outputscale * base_kernel (gpytorch ScaleKernel, Positive constraint)."""

    def __init__(self, base_kernel: _ARDKernel, outputscale_prior=None,
                 outputscale_constraint: GreaterThan | None = None, **kwargs):
        super().__init__()
        self.base_kernel = base_kernel
        c = outputscale_constraint or Positive()
        init = c.initial_value
        if init is None:
            init = max(1.0, c.lower_bound) if c.transform is None else _raw_zero_value(c)
        self._register("outputscale", torch.tensor(float(init), dtype=torch.float64), c,
                       outputscale_prior)

    @property
    def outputscale(self) -> Tensor:
        return self._value("outputscale")

    @outputscale.setter
    def outputscale(self, v) -> None:
        self._set("outputscale", v)

    @property
    def kind(self) -> int:
        return self.base_kernel.kind

    @property
    def lengthscale(self) -> Tensor:
        return self.base_kernel.lengthscale


class GaussianLikelihood(_HyperModule):
    def __init__(self, noise_prior=None, noise_constraint: GreaterThan | None = None, **kwargs):
        super().__init__()
        c = noise_constraint or GreaterThan(1e-4)
        init = c.initial_value
        if init is None:
            init = c.lower_bound if c.transform is None else _raw_zero_value(c)
        self._register("noise", torch.tensor(float(init), dtype=torch.float64), c, noise_prior)

    @property
    def noise(self) -> Tensor:
        return self._value("noise")

    @noise.setter
    def noise(self, v) -> None:
        self._set("noise", v)


class ConstantMean(_HyperModule):
    def __init__(self, **kwargs):
        super().__init__()
        self.raw_constant = torch.nn.Parameter(torch.zeros((), dtype=torch.float64))

    @property
    def constant(self) -> Tensor:
        return self.raw_constant

    @constant.setter
    def constant(self, v) -> None:
        with torch.no_grad():
            self.raw_constant.copy_(torch.as_tensor(v, dtype=torch.float64))


def get_gaussian_likelihood_with_lognormal_prior() -> GaussianLikelihood:
    """models/utils/gpytorch_modules.py:74-97."""
    prior = LogNormalPrior(loc=-4.0, scale=1.0)
    return GaussianLikelihood(
        noise_prior=prior,
        noise_constraint=GreaterThan(MIN_INFERRED_NOISE_LEVEL, transform=None,
                                     initial_value=prior.mode))


def get_covar_module_with_dim_scaled_prior(ard_num_dims: int,
                                           use_rbf_kernel: bool = True) -> _ARDKernel:
    """models/utils/gpytorch_modules.py:100-133."""
    prior = LogNormalPrior(loc=SQRT2 + math.log(ard_num_dims) * 0.5, scale=SQRT3)
    cls = RBFKernel if use_rbf_kernel else MaternKernel
    return cls(ard_num_dims=ard_num_dims, lengthscale_prior=prior,
               lengthscale_constraint=GreaterThan(2.5e-2, transform=None,
                                                  initial_value=prior.mode))


# ----------------------------------------------------------------------------
# Standardize (models/transforms/outcome.py:236-511)
# ----------------------------------------------------------------------------
class Standardize(torch.nn.Module):
    def __init__(self, m: int = 1, min_stdv: float = 1e-8, **kwargs):
        super().__init__()
        if m != 1:
            raise UnsupportedError("only single-output models are supported")
        self._m = m
        self._min_stdv = min_stdv
        self.register_buffer("means", torch.zeros(1, m, dtype=torch.float64))
        self.register_buffer("stdvs", torch.ones(1, m, dtype=torch.float64))

    def forward(self, Y: Tensor, Yvar: Tensor | None = None, X: Tensor | None = None):
        """outcome.py:312-360 in train mode: update the statistics, transform Y."""
        if Y.shape[-2] < 1:
            raise ValueError(f"Can't standardize with no observations. {Y.shape=}.")
        if self.training:
            if Y.shape[-2] == 1:
                stdvs = torch.ones_like(Y[..., :1, :])
            else:
                stdvs = Y.std(dim=-2, keepdim=True)
            stdvs = stdvs.where(stdvs >= self._min_stdv, torch.full_like(stdvs, 1.0))
            self.means = Y.mean(dim=-2, keepdim=True)
            self.stdvs = stdvs
        Y_tf = (Y - self.means) / self.stdvs
        Yvar_tf = Yvar / self.stdvs**2 if Yvar is not None else None
        return Y_tf, Yvar_tf


class _Default:
    pass


DEFAULT = _Default()


# ----------------------------------------------------------------------------
# SingleTaskGP
# ----------------------------------------------------------------------------
class SingleTaskGP(torch.nn.Module):
    """Single-output exact GP (models/gp_regression.py:53-214) whose prediction
    caches, posterior and MLL run in libbotorch_amd."""

    def __init__(self, train_X: Tensor, train_Y: Tensor, train_Yvar: Tensor | None = None,
                 likelihood: GaussianLikelihood | None = None, covar_module=None,
                 mean_module: ConstantMean | None = None, outcome_transform=DEFAULT,
                 input_transform=None):
        super().__init__()
        if train_Yvar is not None:
            raise UnsupportedError("fixed observation noise (train_Yvar) is not supported")
        if input_transform is not None:
            raise UnsupportedError("input transforms are not supported")
        if train_X.dim() != 2 or train_Y.dim() != 2 or train_Y.shape[-1] != 1:
            raise UnsupportedError("expected train_X: n x d and train_Y: n x 1")
        if train_X.shape[0] != train_Y.shape[0]:
            raise InputDataError("train_X and train_Y have different numbers of rows")
        if torch.isnan(train_X).any() or torch.isnan(train_Y).any():
            raise InputDataError("NaN in the training data")
        train_X = train_X.to(torch.float64)
        train_Y = train_Y.to(torch.float64)
        if outcome_transform is DEFAULT:
            outcome_transform = Standardize(m=1)
        if outcome_transform is not None:
            outcome_transform.train()
            train_Y_tf, _ = outcome_transform(train_Y)
            outcome_transform.eval()
            self.outcome_transform = outcome_transform
        else:
            train_Y_tf = train_Y
        d = train_X.shape[-1]
        self.likelihood = likelihood or get_gaussian_likelihood_with_lognormal_prior()
        self.covar_module = covar_module or get_covar_module_with_dim_scaled_prior(d)
        self.mean_module = mean_module or ConstantMean()
        kernel = self.covar_module
        if kernel.lengthscale.numel() != d:
            raise UnsupportedError("the kernel's ard_num_dims must equal train_X.shape[-1]")
        self.train_inputs = (train_X,)
        self.train_targets = train_Y_tf.squeeze(-1).contiguous()
        self._num_outputs = 1
        self._cache = None
        self._cache_key = None
        # Numerics of the prediction caches for ill-conditioned training sets:
        # the factor is chol(K + cholesky_jitter I) and the solves K^-1 (y - c)
        # and K^-1 k(X, X_baseline) take refine_solves refinement steps against K
        # itself.  The defaults (0, 0) are the reference's (plain psd_safe_cholesky).
        self.cholesky_jitter = 0.0
        self.refine_solves = 0
        self.to(train_X.device)

    # -- hyperparameters -----------------------------------------------------
    @property
    def kernel_kind(self) -> int:
        return self.covar_module.kind

    def hyperparameters(self) -> dict:
        k = self.covar_module
        return {
            "lengthscale": k.lengthscale.detach(),
            "outputscale": float(k.outputscale.detach()),
            "noise": float(self.likelihood.noise.detach()),
            "constant": float(self.mean_module.constant.detach()),
        }

    @property
    def y_stats(self) -> tuple[float, float]:
        if hasattr(self, "outcome_transform") and self.outcome_transform is not None:
            return (float(self.outcome_transform.means.reshape(())),
                    float(self.outcome_transform.stdvs.reshape(())))
        return 0.0, 1.0

    @property
    def num_outputs(self) -> int:
        return 1

    @property
    def batch_shape(self) -> torch.Size:
        return torch.Size()

    # -- prediction cache ------------------------------------------------------
    def prediction_cache(self) -> dict:
        """L, L^-1, alpha and the C struct describing the model; rebuilt when a
        hyperparameter changed (like gpytorch's prediction strategy reset)."""
        hp = self.hyperparameters()
        key = (tuple(hp["lengthscale"].tolist()), hp["outputscale"], hp["noise"],
               hp["constant"], self.y_stats, float(self.cholesky_jitter),
               int(self.refine_solves))
        if self._cache is not None and self._cache_key == key:
            return self._cache
        X = self.train_inputs[0]
        dev = X.device
        L = _lib.lib()
        n, d = X.shape
        n_pad = _lib.round_up(n, 128)
        ls = hp["lengthscale"].to(device=dev, dtype=torch.float64).contiguous()
        center = torch.empty(d, dtype=torch.float64, device=dev)
        xs = torch.empty(n_pad, d, dtype=torch.float64, device=dev)
        st = _lib.stream()
        _lib.check(L.bta_scale_inputs(_lib.ptr(X.contiguous()), n, d, None, _lib.ptr(ls),
                                      _lib.ptr(center), _lib.ptr(xs), n_pad, st),
                   "scale_inputs")
        y_mean, y_std = self.y_stats
        m = _lib.GpModel()
        m.kind = self.kernel_kind
        m.d = d
        m.n = n
        m.n_pad = n_pad
        m.train_xs = _lib.ptr(xs)
        m.center = _lib.ptr(center)
        m.lengthscale = _lib.ptr(ls)
        m.outputscale = hp["outputscale"]
        m.noise = hp["noise"]
        m.mean_constant = hp["constant"]
        m.y_mean = y_mean
        m.y_std = y_std
        m.chol_jitter = float(self.cholesky_jitter)
        m.refine_steps = int(self.refine_solves)
        Lmat = torch.empty(n_pad, n_pad, dtype=torch.float64, device=dev)
        Linv = torch.empty(n_pad, n_pad, dtype=torch.float64, device=dev)
        alpha = torch.empty(n_pad, dtype=torch.float64, device=dev)
        ws = _lib.workspace(L.bta_gp_cache_workspace_size(n_pad), dev)
        jit = ctypes.c_double(0.0)
        y = self.train_targets.contiguous()
        _lib.check(L.bta_gp_cache_build(ctypes.byref(m), _lib.ptr(y), _lib.ptr(Lmat),
                                        _lib.ptr(Linv), _lib.ptr(alpha), 6, ctypes.byref(jit),
                                        _lib.ptr(ws), ws.numel(), st), "gp_cache_build")
        m.linv = _lib.ptr(Linv)
        m.alpha = _lib.ptr(alpha)
        self._cache = {"model": m, "xs": xs, "center": center, "lengthscale": ls, "L": Lmat,
                       "Linv": Linv, "alpha": alpha, "n_pad": n_pad, "jitter": jit.value}
        self._cache_key = key
        return self._cache

    # -- posterior (models/gpytorch.py:544-610) -------------------------------
    def posterior(self, X: Tensor, output_indices=None, observation_noise: bool = False,
                  posterior_transform=None) -> "GPyTorchPosterior":
        if posterior_transform is not None:
            raise UnsupportedError("posterior_transform is not supported")
        if output_indices not in (None, [0]):
            raise UnsupportedError("single-output model")
        self.eval()
        X64 = X.to(torch.float64)
        batch_shape = X64.shape[:-2]
        mq, d = X64.shape[-2], X64.shape[-1]
        Xf = X64.reshape(-1, mq, d)
        dev = Xf.device
        if torch.is_grad_enabled() and Xf.requires_grad:
            # differentiable in X like the reference's posterior: the adjoint runs
            # in bta_gp_posterior_backward
            mean, cov = _PosteriorFn.apply(Xf, self)
        else:
            mean, cov = self._posterior_moments(Xf.detach().contiguous())
        if observation_noise is not False:
            # the likelihood's noise (True) or the given per-point noise (a tensor
            # batch x q x 1), in the model's standardized scale, added before the
            # outcome untransform (models/gpytorch.py:514-538)
            _, y_std = self.y_stats
            if torch.is_tensor(observation_noise):
                nz = observation_noise.to(cov).squeeze(-1).expand(*batch_shape, mq)
                cov = cov + (y_std**2) * torch.diag_embed(nz.reshape(-1, mq))
            else:
                cov = cov + (y_std**2) * self.likelihood.noise.detach() * torch.eye(
                    mq, dtype=cov.dtype, device=dev)
        mean = mean.reshape(*batch_shape, mq)
        cov = cov.reshape(*batch_shape, mq, mq)
        return GPyTorchPosterior(mean, cov)

    def _posterior_moments(self, Xf: Tensor):
        """Mean (nb x mq) and covariance (nb x mq x mq), original outcome scale."""
        c = self.prediction_cache()
        L = _lib.lib()
        nb, mq, _ = Xf.shape
        dev = Xf.device
        mean = torch.empty(nb, mq, dtype=torch.float64, device=dev)
        cov = torch.empty(nb, mq, mq, dtype=torch.float64, device=dev)
        ws = _lib.workspace(L.bta_gp_posterior_workspace_size(ctypes.byref(c["model"]), nb, mq),
                            dev)
        _lib.check(L.bta_gp_posterior(ctypes.byref(c["model"]), _lib.ptr(Xf), nb, mq,
                                      _lib.ptr(mean), _lib.ptr(cov), None, _lib.ptr(ws),
                                      ws.numel(), _lib.stream()), "gp_posterior")
        return mean, cov

    def _posterior_backward(self, Xf: Tensor, dmean: Tensor | None, dcov: Tensor | None):
        """d X (nb x mq x d) of <dmean, mean> + <dcov, covariance>."""
        c = self.prediction_cache()
        L = _lib.lib()
        nb, mq, d = Xf.shape
        dev = Xf.device
        dX = torch.empty(nb, mq, d, dtype=torch.float64, device=dev)
        if dmean is None and dcov is None:
            return dX.zero_()
        ws = _lib.workspace(L.bta_gp_posterior_backward_workspace_size(
            ctypes.byref(c["model"]), nb, mq), dev)
        dm = None if dmean is None else dmean.to(torch.float64).contiguous()
        dc = None if dcov is None else dcov.to(torch.float64).contiguous()
        _lib.check(L.bta_gp_posterior_backward(ctypes.byref(c["model"]), _lib.ptr(Xf), nb, mq,
                                               _lib.ptr(dm), _lib.ptr(dc), _lib.ptr(dX),
                                               _lib.ptr(ws), ws.numel(), _lib.stream()),
                   "gp_posterior_backward")
        return dX

    def forward(self, X: Tensor):
        raise UnsupportedError("use posterior(); the prior MVN is not materialised")


class _PosteriorFn(torch.autograd.Function):
    """Posterior moments with the adjoint of bta_gp_posterior_backward (the
    reference keeps SingleTaskGP.posterior differentiable in X,
    models/gpytorch.py:544-610)."""

    @staticmethod
    def forward(ctx, Xf: Tensor, model: "SingleTaskGP"):
        Xc = Xf.detach().contiguous()
        mean, cov = model._posterior_moments(Xc)
        ctx.model = model
        ctx.save_for_backward(Xc)
        return mean, cov

    @staticmethod
    def backward(ctx, dmean: Tensor | None, dcov: Tensor | None):
        (Xc,) = ctx.saved_tensors
        return ctx.model._posterior_backward(Xc, dmean, dcov), None


# ----------------------------------------------------------------------------
# posterior (posteriors/gpytorch.py:35-180) for a single-output MVN
# ----------------------------------------------------------------------------
def psd_safe_cholesky(A: Tensor, max_tries: int = 6, jitter: float = 1e-8) -> Tensor:
    """Cholesky of ... x k x k matrices on the GPU with jitter escalation
    (linear_operator psd_safe_cholesky semantics)."""
    L = _lib.lib()
    k = A.shape[-1]
    batch = A.shape[:-2]
    Af = A.to(torch.float64).reshape(-1, k, k).contiguous().clone()
    nb = Af.shape[0]
    st = _lib.stream()
    if k <= 64:
        status = torch.zeros(nb, dtype=torch.int32, device=A.device)
        _lib.check(L.bta_chol_small_batched(_lib.ptr(Af), k, k, k * k, nb, max_tries, jitter,
                                            _lib.ptr(status), st), "chol_small_batched")
        flags = [bool(((status & f) != 0).any()) for f in (_lib.STATUS_NAN, _lib.STATUS_NOT_PSD)]
        if flags[0]:
            raise NanError("cholesky input contains NaNs")
        if flags[1]:
            raise NotPSDError("Matrix not positive definite after repeatedly adding jitter")
        return Af.reshape(*batch, k, k)
    kp = _lib.round_up(k, 128)
    out = torch.empty(nb, k, k, dtype=torch.float64, device=A.device)
    work = torch.zeros(kp, kp, dtype=torch.float64, device=A.device)
    ws = _lib.workspace(L.bta_potrf_workspace_size(kp), A.device)
    jit = ctypes.c_double(0.0)
    for i in range(nb):
        work.zero_()
        work[:k, :k] = Af[i]
        if kp > k:
            idx = torch.arange(k, kp, device=A.device)
            work[idx, idx] = 1.0
        _lib.check(L.bta_potrf(_lib.ptr(work), kp, kp, None, 0, max_tries, jitter,
                               ctypes.byref(jit), _lib.ptr(ws), ws.numel(), st), "potrf")
        out[i] = work[:k, :k]
    return out.reshape(*batch, k, k)


class GPyTorchPosterior:
    """Posterior of a single-output GP: mean ``batch x m``, covariance
    ``batch x m x m`` (original outcome scale)."""

    def __init__(self, mean: Tensor, covariance: Tensor):
        self._mean = mean
        self._covar = covariance
        self._scale_tril = None
        self._is_mt = False

    @property
    def device(self):
        return self._mean.device

    @property
    def dtype(self):
        return self._mean.dtype

    @property
    def batch_shape(self) -> torch.Size:
        return self._mean.shape[:-1]

    @property
    def mean(self) -> Tensor:
        return self._mean.unsqueeze(-1)

    @property
    def variance(self) -> Tensor:
        return torch.diagonal(self._covar, dim1=-2, dim2=-1).unsqueeze(-1)

    @property
    def covariance_matrix(self) -> Tensor:
        return self._covar

    @property
    def base_sample_shape(self) -> torch.Size:
        return self._mean.shape

    @property
    def batch_range(self) -> tuple[int, int]:
        return (0, -1)

    def _extended_shape(self, sample_shape: torch.Size = torch.Size()) -> torch.Size:
        return sample_shape + self._mean.shape + torch.Size([1])

    @property
    def scale_tril(self) -> Tensor:
        if self._scale_tril is None:
            if torch.is_grad_enabled() and self._covar.requires_grad:
                self._scale_tril = _CholFn.apply(self._covar)
            else:
                self._scale_tril = psd_safe_cholesky(self._covar)
        return self._scale_tril

    def rsample_from_base_samples(self, sample_shape: torch.Size, base_samples: Tensor) -> Tensor:
        """mean + L z (posteriors/gpytorch.py:86-127, gpytorch MVN.rsample)."""
        if base_samples.shape[: len(sample_shape)] != sample_shape:
            raise RuntimeError("`sample_shape` disagrees with shape of `base_samples`.")
        Lt = self.scale_tril
        m = self._mean.shape[-1]
        z = base_samples.to(self._mean).expand(*sample_shape, *self._mean.shape)
        S = int(torch.Size(sample_shape).numel())
        nb = int(self._mean[..., 0].numel())
        zf = z.reshape(S, nb, m).permute(1, 0, 2).contiguous()  # nb x S x m
        if torch.is_grad_enabled() and Lt.requires_grad:
            out = _LowerMatmulFn.apply(Lt.reshape(nb, m, m), zf)
        else:
            out = _batched_lower_matmul(Lt.reshape(nb, m, m), zf)  # nb x S x m: z L^T
        out = out + self._mean.reshape(nb, 1, m)
        return out.permute(1, 0, 2).reshape(*sample_shape, *self._mean.shape, 1)

    def rsample(self, sample_shape: torch.Size | None = None) -> Tensor:
        sample_shape = sample_shape or torch.Size([1])
        z = torch.randn(*sample_shape, *self._mean.shape, dtype=self.dtype, device=self.device)
        return self.rsample_from_base_samples(sample_shape, z)


def _batched_lower_matmul(Lt: Tensor, Z: Tensor) -> Tensor:
    """Z L^T for L: nb x m x m (lower), Z: nb x S x m, on the MFMA GEMM."""
    lib = _lib.lib()
    nb, S, m = Z.shape
    Sp, mp = _lib.round_up(S, 64), _lib.round_up(m, 64)
    Zp = torch.zeros(nb, Sp, mp, dtype=torch.float64, device=Z.device)
    Zp[:, :S, :m] = Z
    Lp = torch.zeros(nb, mp, mp, dtype=torch.float64, device=Z.device)
    Lp[:, :m, :m] = Lt
    Cp = torch.empty(nb, Sp, mp, dtype=torch.float64, device=Z.device)
    _lib.check(lib.bta_gemm(_lib.GEMM_B_LOWER, Sp, mp, mp, 1.0, _lib.ptr(Zp), mp, _lib.ptr(Lp),
                            mp, 0.0, None, 0, _lib.ptr(Cp), mp, nb, Sp * mp, mp * mp, 0,
                            Sp * mp, _lib.stream()), "gemm")
    return Cp[:, :S, :m]


class _CholFn(torch.autograd.Function):
    """psd_safe_cholesky with the Cholesky adjoint on the GPU
    (bta_chol_small_backward; k <= 64)."""

    @staticmethod
    def forward(ctx, A: Tensor):
        Lt = psd_safe_cholesky(A.detach())
        ctx.save_for_backward(Lt)
        return Lt

    @staticmethod
    def backward(ctx, dL: Tensor):
        (Lt,) = ctx.saved_tensors
        k = Lt.shape[-1]
        if k > 64:
            raise UnsupportedError("the Cholesky adjoint is implemented for k <= 64")
        Lf = Lt.reshape(-1, k, k).contiguous()
        dLf = dL.to(torch.float64).reshape(-1, k, k).contiguous()
        dA = torch.empty_like(Lf)
        _lib.check(_lib.lib().bta_chol_small_backward(_lib.ptr(Lf), _lib.ptr(dLf), k, k * k,
                                                      Lf.shape[0], _lib.ptr(dA), _lib.stream()),
                   "chol_small_backward")
        return dA.reshape(Lt.shape)


class _LowerMatmulFn(torch.autograd.Function):
    """Z L^T for lower L (nb x m x m) and Z (nb x S x m) on the MFMA GEMM, with
    dL = tril(dOut^T Z) on the same GEMM."""

    @staticmethod
    def forward(ctx, Lt: Tensor, Z: Tensor):
        ctx.save_for_backward(Z)
        return _batched_lower_matmul(Lt.detach(), Z)

    @staticmethod
    def backward(ctx, dOut: Tensor):
        (Z,) = ctx.saved_tensors
        lib = _lib.lib()
        nb, S, m = Z.shape
        Sp, mp = _lib.round_up(S, 64), _lib.round_up(m, 64)
        Gp = torch.zeros(nb, Sp, mp, dtype=torch.float64, device=Z.device)
        Gp[:, :S, :m] = dOut
        Zp = torch.zeros(nb, Sp, mp, dtype=torch.float64, device=Z.device)
        Zp[:, :S, :m] = Z
        Cp = torch.empty(nb, mp, mp, dtype=torch.float64, device=Z.device)
        # C[i][j] = sum_s dOut[s][i] Z[s][j]  (A stored K x M, B stored K x N)
        _lib.check(lib.bta_gemm(_lib.GEMM_A_KM | _lib.GEMM_B_KN, mp, mp, Sp, 1.0, _lib.ptr(Gp),
                                mp, _lib.ptr(Zp), mp, 0.0, None, 0, _lib.ptr(Cp), mp, nb,
                                Sp * mp, Sp * mp, 0, mp * mp, _lib.stream()), "gemm")
        return torch.tril(Cp[:, :m, :m]), None
